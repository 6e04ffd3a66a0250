// sgeo_batch.hip -- multi-trial (multi-RHS) kernels of the mixed-precision
// EBE-PCG: B strength-reduction trials share one mesh and one elastic
// stiffness (the viscoplastic method never changes K, src/slope.F90:288), so
// every CG iteration streams K once for all B columns.
//
// Vectors are interleaved [n1][B] (column fastest): an element gathers 81
// contiguous B-vectors and every update is unit-stride.  Inner CG vectors
// and the inner element product are FP32; dot products, the refinement
// residual f - K x (FP64 operator), x and all plastic state are FP64.
// Every reduction is two-phase (per-block partials, then one block per
// column) and combines in a fixed order, so a column's arithmetic does not
// depend on B or on the other columns: a trial gives the same bits alone and
// in any batch.
#pragma clang fp contract(off)

#include "sgeo_device.h"
#include "sgeo_internal.h"

namespace sgeo {

// --------------------------------------------------------------------------
// batched element product: Y[e][r][j] = sum_c K(r,c) P[g(e,c)][j]
// --------------------------------------------------------------------------
// One element per block.  The packed upper triangle of K_e (KT) and the
// gathered P_e (81 x B, VT) are staged in LDS; thread t owns column
// j = t % B and rows r = t / B + k * (nthreads / B).  Rows sum c = 0..80 in
// order in AT precision.  Columns with cmask[j] != want read zero and write
// nothing.
// Columns col0 .. col0+B-1 of arrays with column stride Bs are processed.
template <typename KT, typename VT, typename AT>
__global__ __launch_bounds__(256) void k_bmv(int32_t nelmt, int B, int Bs,
                                             int col0,
                                             const KT* __restrict__ KP,
                                             const int32_t* __restrict__ gdof,
                                             const VT* __restrict__ P,
                                             VT* __restrict__ Y,
                                             const int32_t* __restrict__ cst,
                                             int want) {
  extern __shared__ __align__(16) unsigned char smem[];
  KT* u = reinterpret_cast<KT*>(smem);
  VT* pe = reinterpret_cast<VT*>(smem + SGEO_KPACK * sizeof(KT));
  const int64_t e = blockIdx.x;
  const int t = threadIdx.x, nt = blockDim.x;
  const KT* src = KP + e * SGEO_KPACK;
  for (int q = t; q < SGEO_KPACK; q += nt) u[q] = src[q];
  for (int q = t; q < 81 * B; q += nt) {
    const int c = q / B, j = q - c * B;
    pe[q] = (cst[col0 + j] == want) ? P[(int64_t)gdof[e * 81 + c] * Bs + col0 + j]
                                    : (VT)0;
  }
  __syncthreads();
  const int j = t % B;
  const int nrg = nt / B, rg = t / B;
  if (rg >= nrg || cst[col0 + j] != want) return;
  constexpr int MAXR = 81;
  AT acc[MAXR / 4 + 1];
  const int nr = (81 - rg + nrg - 1) / nrg;  // rows rg, rg+nrg, ...
#pragma unroll
  for (int k = 0; k < MAXR / 4 + 1; k++) acc[k] = (AT)0;
  if (nr <= MAXR / 4 + 1) {
    for (int c = 0; c < 81; c++) {
      const AT pv = (AT)pe[c * B + j];
      const int cb = c * (c + 1) / 2;
#pragma unroll
      for (int k = 0; k < MAXR / 4 + 1; k++) {
        const int r = rg + k * nrg;
        if (k < nr) {
          const int idx = (r <= c) ? (cb + r) : (r * (r + 1) / 2 + c);
          acc[k] = acc[k] + (AT)u[idx] * pv;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < MAXR / 4 + 1; k++) {
      const int r = rg + k * nrg;
      if (k < nr) Y[(e * 81 + r) * Bs + col0 + j] = (VT)acc[k];
    }
  } else {
    // fewer than 4 row groups (B > 64 / nt): rows one by one
    for (int r = rg; r < 81; r += nrg) {
      AT a = (AT)0;
      for (int c = 0; c < 81; c++) {
        const int idx = (r <= c) ? (c * (c + 1) / 2 + r) : (r * (r + 1) / 2 + c);
        a = a + (AT)u[idx] * (AT)pe[c * B + j];
      }
      Y[(e * 81 + r) * Bs + col0 + j] = (VT)a;
    }
  }
}

// --------------------------------------------------------------------------
// batched vector kernels.  Equations are split into fixed chunks of
// BCHUNK; thread t owns column j = t % B and chunk blockIdx*(256/B) + t/B,
// and walks its chunk in order, so every per-(chunk, column) partial is a
// sequential sum whose value does not depend on B.  The one-block-per-column
// finalize then combines the chunk partials in a fixed order.
// Partials: part[(kind*nch + chunk)*B + j].
// --------------------------------------------------------------------------
#define BCHUNK 64

struct ChunkIdx {
  int j;
  int64_t chunk, g0, g1;
  bool valid;
};

__device__ __forceinline__ ChunkIdx chunk_of(int64_t n1, int B) {
  ChunkIdx c;
  const int t = threadIdx.x;
  c.j = t % B;
  c.chunk = (int64_t)blockIdx.x * (blockDim.x / B) + t / B;
  c.g0 = c.chunk * BCHUNK;
  c.g1 = c.g0 + BCHUNK < n1 ? c.g0 + BCHUNK : n1;
  c.valid = c.g0 < n1;
  return c;
}

__device__ __forceinline__ int64_t nchunks(int64_t n1) {
  return (n1 + BCHUNK - 1) / BCHUNK;
}

// kp[g][j] = sum over slots of Y (element order, FP32); partial p.kp (FP64)
// and max|p| for columns in state INNER
__global__ __launch_bounds__(256) void k_bassemble_dot(
    int64_t n1, int B, const int32_t* __restrict__ rowptr,
    const int32_t* __restrict__ slots, const float* __restrict__ Y,
    const float* __restrict__ p, float* __restrict__ kp,
    double* __restrict__ part, const int32_t* __restrict__ cst) {
  const ChunkIdx c = chunk_of(n1, B);
  if (!c.valid || cst[c.j] != BCOL_INNER) return;
  double s = 0.0, m = 0.0;
  for (int64_t g = c.g0; g < c.g1; g++) {
    float acc = 0.0f;
    if (g > 0)
      for (int32_t q = rowptr[g]; q < rowptr[g + 1]; q++)
        acc = acc + Y[(int64_t)slots[q] * B + c.j];
    kp[g * B + c.j] = acc;
    const double pg = (double)p[g * B + c.j];
    s = s + pg * (double)acc;
    m = fmax(m, fabs(pg));
  }
  const int64_t nch = nchunks(n1);
  part[(0 * nch + c.chunk) * B + c.j] = s;
  part[(1 * nch + c.chunk) * B + c.j] = m;
}

// d += alpha p; r -= alpha kp; z = dinv r; partial max|d| and r.z (FP64)
__global__ __launch_bounds__(256) void k_bupdate(
    int64_t n1, int B, const float* __restrict__ p, const float* __restrict__ kp,
    const float* __restrict__ dinv, float* __restrict__ d, float* __restrict__ r,
    float* __restrict__ z, double* __restrict__ part,
    const int32_t* __restrict__ cst, const BCol* __restrict__ bc) {
  const ChunkIdx c = chunk_of(n1, B);
  if (!c.valid || cst[c.j] != BCOL_INNER) return;
  const float alpha = (float)bc[c.j].alpha;
  double md = 0.0, s = 0.0;
  for (int64_t g = c.g0; g < c.g1; g++) {
    const int64_t q = g * B + c.j;
    const float dg = d[q] + alpha * p[q];
    d[q] = dg;
    md = fmax(md, fabs((double)dg));
    const float rg = r[q] - alpha * kp[q];
    const float zg = dinv[g] * rg;
    r[q] = rg;
    z[q] = zg;
    s = s + (double)rg * (double)zg;
  }
  const int64_t nch = nchunks(n1);
  part[(0 * nch + c.chunk) * B + c.j] = md;
  part[(1 * nch + c.chunk) * B + c.j] = s;
}

// p = z + beta p
__global__ __launch_bounds__(256) void k_bpupdate(int64_t n1, int B,
                                                  const float* __restrict__ z,
                                                  float* __restrict__ p,
                                                  const int32_t* __restrict__ cst,
                                                  const BCol* __restrict__ bc) {
  const int64_t n = n1 * B;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(q % B);
    if (cst[j] != BCOL_INNER) continue;
    p[q] = z[q] + (float)bc[j].beta * p[q];
  }
}

// refinement residual and inner start, columns in NEED_RESID:
// r64 = load - sum Y64 (element order, FP64); r = (float) r64;
// z = dinv r; p = z; d = 0; partial r.z and max|r64|
__global__ __launch_bounds__(256) void k_bresid(
    int64_t n1, int B, const int32_t* __restrict__ rowptr,
    const int32_t* __restrict__ slots, const double* __restrict__ Y64,
    const double* __restrict__ load, const float* __restrict__ dinv,
    float* __restrict__ r, float* __restrict__ z, float* __restrict__ p,
    float* __restrict__ d, double* __restrict__ part,
    const int32_t* __restrict__ cst) {
  const ChunkIdx c = chunk_of(n1, B);
  if (!c.valid || cst[c.j] != BCOL_NEED_RESID) return;
  double s = 0.0, m = 0.0;
  for (int64_t g = c.g0; g < c.g1; g++) {
    const int64_t q = g * B + c.j;
    double kx = 0.0;
    if (g > 0)
      for (int32_t w = rowptr[g]; w < rowptr[g + 1]; w++)
        kx = kx + Y64[(int64_t)slots[w] * B + c.j];
    const double r64 = (g == 0) ? 0.0 : load[q] - kx;
    const float rg = (float)r64;
    const float zg = dinv[g] * rg;
    r[q] = rg;
    z[q] = zg;
    p[q] = zg;
    d[q] = 0.0f;
    s = s + (double)rg * (double)zg;
    m = fmax(m, fabs(r64));
  }
  const int64_t nch = nchunks(n1);
  part[(0 * nch + c.chunk) * B + c.j] = s;
  part[(1 * nch + c.chunk) * B + c.j] = m;
}

// x += d for columns in INNER_DONE; partial max|d| and max|x| (after update)
__global__ __launch_bounds__(256) void k_bxupdate(
    int64_t n1, int B, const float* __restrict__ d, double* __restrict__ x,
    double* __restrict__ part, const int32_t* __restrict__ cst) {
  const ChunkIdx c = chunk_of(n1, B);
  if (!c.valid || cst[c.j] != BCOL_INNER_DONE) return;
  double md = 0.0, mx = 0.0;
  for (int64_t g = c.g0; g < c.g1; g++) {
    const int64_t q = g * B + c.j;
    const double dg = (g == 0) ? 0.0 : (double)d[q];
    const double xg = x[q] + dg;
    x[q] = xg;
    md = fmax(md, fabs(dg));
    mx = fmax(mx, fabs(xg));
  }
  const int64_t nch = nchunks(n1);
  part[(0 * nch + c.chunk) * B + c.j] = md;
  part[(1 * nch + c.chunk) * B + c.j] = mx;
}

// one block per column: combine the chunk partials of column j in a fixed
// order and advance the column's state machine
__global__ __launch_bounds__(256) void k_bfinalize(int64_t nch, int B,
                                                   const double* __restrict__ part,
                                                   int32_t* __restrict__ cst,
                                                   BCol* __restrict__ bc,
                                                   int what) {
  __shared__ double red[4];
  const int j = blockIdx.x;
  const int st = cst[j];
  const bool mine = (what == BFIN_RZ && st == BCOL_NEED_RESID) ||
                    ((what == BFIN_ALPHA || what == BFIN_INNER) && st == BCOL_INNER) ||
                    (what == BFIN_OUTER && st == BCOL_INNER_DONE);
  if (!mine) return;  // uniform per block
  const bool max0 = (what == BFIN_INNER || what == BFIN_OUTER);
  const bool max1 = (what != BFIN_INNER);
  double v0 = 0.0, v1 = 0.0;
  for (int64_t b = threadIdx.x; b < nch; b += 256) {
    const double a0 = part[(0 * nch + b) * B + j];
    const double a1 = part[(1 * nch + b) * B + j];
    v0 = max0 ? fmax(v0, a0) : v0 + a0;
    v1 = max1 ? fmax(v1, a1) : v1 + a1;
  }
  v0 = max0 ? block_max<4>(v0, red) : block_sum<4>(v0, red);
  v1 = max1 ? block_max<4>(v1, red) : block_sum<4>(v1, red);
  if (threadIdx.x != 0) return;
  BCol& c = bc[j];
  switch (what) {
    case BFIN_RZ:  // v0 = r.z, v1 = max|r64|
      c.rz = v0;
      c.inner_it = 0;
      c.outer_it += 1;
      cst[j] = (v1 == 0.0) ? BCOL_INNER_DONE : BCOL_INNER;  // exact: d = 0
      break;
    case BFIN_ALPHA:  // v0 = p.kp, v1 = max|p|
      c.maxp = v1;
      c.alpha = c.rz / v0;
      break;
    case BFIN_INNER: {  // v0 = max|d|, v1 = r.z
      c.inner_it += 1;
      c.cg_total += 1;
      const bool stop = fabs(c.alpha) * c.maxp / v0 <= c.inner_tol;
      if (stop || c.inner_it >= c.inner_max || c.cg_total >= c.cg_maxiter) {
        cst[j] = BCOL_INNER_DONE;
      } else {
        c.beta = v1 / c.rz;
        c.rz = v1;
      }
      break;
    }
    case BFIN_OUTER:  // v0 = max|d|, v1 = max|x|
      c.maxd = v0;
      c.maxx = v1;
      if (v0 / v1 <= c.tol) cst[j] = BCOL_DONE;
      else if (c.cg_total >= c.cg_maxiter || c.outer_it >= c.outer_max)
        cst[j] = BCOL_FAIL;
      else cst[j] = BCOL_NEED_RESID;
      break;
  }
}

// load[g][j] = ext[g] + body[g][j] (load(0)=0) for active columns
__global__ void k_bload(int64_t n1, int B, const double* __restrict__ ext,
                        const double* __restrict__ body, double* __restrict__ load,
                        const int32_t* __restrict__ active) {
  const int64_t n = n1 * B;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = q / B;
    const int j = (int)(q - g * B);
    if (!active[j]) continue;
    load[q] = (g == 0) ? 0.0 : ext[g] + body[q];
  }
}

// uerr per column: x(0)=0, partial max|x-oldx| and max|x|, oldx = x
__global__ __launch_bounds__(256) void k_buerr(int64_t n1, int B,
                                               double* __restrict__ x,
                                               double* __restrict__ oldx,
                                               double* __restrict__ part,
                                               const int32_t* __restrict__ active) {
  const ChunkIdx c = chunk_of(n1, B);
  if (!c.valid || !active[c.j]) return;
  double md = 0.0, mx = 0.0;
  for (int64_t g = c.g0; g < c.g1; g++) {
    const int64_t q = g * B + c.j;
    double xg = x[q];
    if (g == 0) { xg = 0.0; x[q] = 0.0; }
    md = fmax(md, fabs(xg - oldx[q]));
    mx = fmax(mx, fabs(xg));
    oldx[q] = xg;
  }
  const int64_t nch = nchunks(n1);
  part[(0 * nch + c.chunk) * B + c.j] = md;
  part[(1 * nch + c.chunk) * B + c.j] = mx;
}

__global__ __launch_bounds__(256) void k_buerr_finalize(
    int64_t nch, int B, const double* __restrict__ part,
    PlasticScalars* __restrict__ ps, const int32_t* __restrict__ active,
    double nl_tol) {
  __shared__ double red[4];
  const int j = blockIdx.x;
  if (!active[j]) return;
  double vd = 0.0, vx = 0.0;
  for (int64_t b = threadIdx.x; b < nch; b += 256) {
    vd = fmax(vd, part[(0 * nch + b) * B + j]);
    vx = fmax(vx, part[(1 * nch + b) * B + j]);
  }
  vd = block_max<4>(vd, red);
  vx = block_max<4>(vx, red);
  if (threadIdx.x == 0) {
    const double uerr = vd / vx;
    ps[j].uerr = uerr;
    ps[j].xmax = vx;
    ps[j].nl_isconv = uerr <= nl_tol ? 1 : 0;
    ps[j].fmax_bits = 0ull;  // fmax = zero (src/slope.F90:286)
    ps[j].n_plastic = 0;
  }
}

// body[g][j] += its element body loads (element order) unless the column's
// trial converged or is on its last iteration (src/slope.F90:380-384)
__global__ void k_bbody_assemble(int64_t n1, int B, int32_t nel,
                                 const int32_t* __restrict__ rowptr,
                                 const int32_t* __restrict__ slots,
                                 const double* __restrict__ BL,
                                 double* __restrict__ body,
                                 const int32_t* __restrict__ active,
                                 const int32_t* __restrict__ is_last,
                                 const PlasticScalars* __restrict__ ps) {
  const int64_t n = n1 * B;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = q / B;
    const int j = (int)(q - g * B);
    if (!active[j] || is_last[j] || ps[j].nl_isconv) continue;
    if (g == 0) { body[q] = 0.0; continue; }
    double acc = body[q];
    const double* blj = BL + (int64_t)j * nel * 81;
    for (int32_t w = rowptr[g]; w < rowptr[g + 1]; w++) acc = acc + blj[slots[w]];
    body[q] = acc;
  }
}

// column j of an interleaved [n1][B] array: set from src (contiguous n1,
// NULL = zero) or copy out to dst (contiguous)
__global__ void k_col_io(int64_t n1, int B, int j, double* __restrict__ a,
                         const double* __restrict__ src, double* __restrict__ dst) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n1;
       g += (int64_t)gridDim.x * blockDim.x) {
    if (dst) dst[g] = a[g * B + j];
    else a[g * B + j] = src ? src[g] : 0.0;
  }
}

// FP32 copy of an FP64 vector
__global__ void k_to_f32(int64_t n, const double* __restrict__ a,
                         float* __restrict__ b) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
       g += (int64_t)gridDim.x * blockDim.x)
    b[g] = (float)a[g];
}

// --------------------------------------------------------------------------
// launch wrappers
// --------------------------------------------------------------------------
static inline unsigned grid_stride(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

static inline int64_t nchunks_host(int64_t n1) {
  return (n1 + BCHUNK - 1) / BCHUNK;
}

int64_t batch_partials(int64_t n1) { return 2 * ((n1 + BCHUNK - 1) / BCHUNK); }

static inline unsigned chunk_blocks(int64_t n1, int B) {
  const int64_t per = 256 / B;  // chunks per block
  return (unsigned)((nchunks_host(n1) + per - 1) / per);
}

static inline int bmv_threads(int B) {
  int t = 81 * B;
  if (t > 256) t = 256;
  t = (t + 63) / 64 * 64;
  return t < 64 ? 64 : t;
}

void launch_bmv_f32(hipStream_t s, int32_t nelmt, int B, const float* KP,
                    const int32_t* gdof, const float* P, float* Y,
                    const int32_t* cst, int want) {
  const size_t lds = SGEO_KPACK * sizeof(float) + (size_t)81 * B * sizeof(float);
  hipLaunchKernelGGL((k_bmv<float, float, float>), dim3(nelmt), dim3(bmv_threads(B)),
                     lds, s, nelmt, B, B, 0, KP, gdof, P, Y, cst, want);
}

// FP64: columns in chunks of at most 32 (LDS <= 48 KB per block)
void launch_bmv_f64(hipStream_t s, int32_t nelmt, int B, const double* KP,
                    const int32_t* gdof, const double* X, double* Y,
                    const int32_t* cst, int want) {
  const int Bc = B < 32 ? B : 32;
  const size_t lds = SGEO_KPACK * sizeof(double) + (size_t)81 * Bc * sizeof(double);
  for (int c0 = 0; c0 < B; c0 += Bc)
    hipLaunchKernelGGL((k_bmv<double, double, double>), dim3(nelmt),
                       dim3(bmv_threads(Bc)), lds, s, nelmt, Bc, B, c0, KP, gdof,
                       X, Y, cst, want);
}

void launch_bassemble_dot(hipStream_t s, int64_t n1, int B, const int32_t* rowptr,
                          const int32_t* slots, const float* Y, const float* p,
                          float* kp, double* part, const int32_t* cst) {
  hipLaunchKernelGGL(k_bassemble_dot, dim3(chunk_blocks(n1, B)), dim3(256), 0, s,
                     n1, B, rowptr, slots, Y, p, kp, part, cst);
}

void launch_bupdate(hipStream_t s, int64_t n1, int B, const float* p,
                    const float* kp, const float* dinv, float* d, float* r,
                    float* z, double* part, const int32_t* cst, const BCol* bc) {
  hipLaunchKernelGGL(k_bupdate, dim3(chunk_blocks(n1, B)), dim3(256), 0, s, n1, B,
                     p, kp, dinv, d, r, z, part, cst, bc);
}

void launch_bpupdate(hipStream_t s, int64_t n1, int B, const float* z, float* p,
                     const int32_t* cst, const BCol* bc) {
  hipLaunchKernelGGL(k_bpupdate, dim3(grid_stride(n1 * B)), dim3(256), 0, s, n1,
                     B, z, p, cst, bc);
}

void launch_bresid(hipStream_t s, int64_t n1, int B, const int32_t* rowptr,
                   const int32_t* slots, const double* Y64, const double* load,
                   const float* dinv, float* r, float* z, float* p, float* d,
                   double* part, const int32_t* cst) {
  hipLaunchKernelGGL(k_bresid, dim3(chunk_blocks(n1, B)), dim3(256), 0, s, n1, B,
                     rowptr, slots, Y64, load, dinv, r, z, p, d, part, cst);
}

void launch_bxupdate(hipStream_t s, int64_t n1, int B, const float* d, double* x,
                     double* part, const int32_t* cst) {
  hipLaunchKernelGGL(k_bxupdate, dim3(chunk_blocks(n1, B)), dim3(256), 0, s, n1,
                     B, d, x, part, cst);
}

void launch_bfinalize(hipStream_t s, int64_t n1, int B, const double* part,
                      int32_t* cst, BCol* bc, int what) {
  hipLaunchKernelGGL(k_bfinalize, dim3(B), dim3(256), 0, s, nchunks_host(n1), B,
                     part, cst, bc, what);
}

void launch_bload(hipStream_t s, int64_t n1, int B, const double* ext,
                  const double* body, double* load, const int32_t* active) {
  hipLaunchKernelGGL(k_bload, dim3(grid_stride(n1 * B)), dim3(256), 0, s, n1, B,
                     ext, body, load, active);
}

void launch_buerr(hipStream_t s, int64_t n1, int B, double* x, double* oldx,
                  double* part, PlasticScalars* ps, const int32_t* active,
                  double nl_tol) {
  hipLaunchKernelGGL(k_buerr, dim3(chunk_blocks(n1, B)), dim3(256), 0, s, n1, B,
                     x, oldx, part, active);
  hipLaunchKernelGGL(k_buerr_finalize, dim3(B), dim3(256), 0, s, nchunks_host(n1),
                     B, part, ps, active, nl_tol);
}

void launch_col_io(hipStream_t s, int64_t n1, int B, int j, double* a,
                   const double* src, double* dst) {
  hipLaunchKernelGGL(k_col_io, dim3(grid_stride(n1)), dim3(256), 0, s, n1, B, j,
                     a, src, dst);
}

void launch_to_f32(hipStream_t s, int64_t n, const double* a, float* b) {
  hipLaunchKernelGGL(k_to_f32, dim3(grid_stride(n)), dim3(256), 0, s, n, a, b);
}

void launch_bbody_assemble(hipStream_t s, int64_t n1, int B, int32_t nel,
                           const int32_t* rowptr, const int32_t* slots,
                           const double* BL, double* body, const int32_t* active,
                           const int32_t* is_last, const PlasticScalars* ps) {
  hipLaunchKernelGGL(k_bbody_assemble, dim3(grid_stride(n1 * B)), dim3(256), 0, s,
                     n1, B, nel, rowptr, slots, BL, body, active, is_last, ps);
}

}  // namespace sgeo
