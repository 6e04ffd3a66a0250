// sgeo_pcg.hip -- EBE matrix-vector product, assembly and the Jacobi-PCG
// vector kernels (src/solver.f90:14-79) for gfx950.
//
// Reductions are two-phase: every block of a vector kernel writes its
// partial sums/maxima to a scratch array, then a one-block finalize kernel
// combines them in index order and updates the device-resident scalars.
// The kernel boundary provides the visibility; no device-scope fences or
// atomics are issued per block (on a multi-XCD MI355X a device-scope
// release writes back the L2, which made a last-block-done pattern cost an
// order of magnitude more than the streaming itself).  Partials are
// combined in a fixed order, so results are deterministic run to run.
//
// Contraction is off: element products and vector updates round each
// operation as the reference does.
#pragma clang fp contract(off)

#include <type_traits>

#include "sgeo_device.h"
#include "sgeo_internal.h"

namespace sgeo {

static inline unsigned nblk(int64_t n, int per) {
  return (unsigned)((n + per - 1) / per);
}

static inline unsigned grid_stride_blocks(int64_t n) {
  const int64_t b = (n + SGEO_VEC_THREADS - 1) / SGEO_VEC_THREADS;
  return (unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

// --------------------------------------------------------------------------
// EBE product, element part: Y[e] = K_e p_e
// --------------------------------------------------------------------------
// Three elements per 256-thread block: thread t owns row r = t % 81 of
// element 3*blockIdx + t / 81 (threads 243..255 idle).  Column c of the
// three column-major K_e is read by the block as three contiguous 648-byte
// runs, so every wave-wide load is (nearly) full width.  p_e is gathered
// into LDS once; row r sums c = 0..80 in order, as matmul(km, p(egdof))
// (src/solver.f90:57).
// skip_mode: 0 none, 1 return when sc->done, 2 return when !sc->u_nonzero.
template <typename KT>
__global__ __launch_bounds__(256) void k_ebe_elem(
    int32_t nelmt, const KT* __restrict__ K, const int32_t* __restrict__ gdof,
    const double* __restrict__ p, double* __restrict__ Y,
    const PcgScalars* __restrict__ sc, int skip_mode) {
  if (skip_mode == 1 && sc->done) return;
  if (skip_mode == 2 && !sc->u_nonzero) return;
  __shared__ double pe[3][81];
  const int t = threadIdx.x;
  const int el = t / 81, r = t - 81 * el;
  const int64_t e = (int64_t)blockIdx.x * 3 + el;
  const bool valid = (t < 243) && (e < nelmt);
  if (valid) pe[el][r] = p[gdof[e * 81 + r]];
  __syncthreads();
  if (!valid) return;
  const KT* ke = K + e * 6561 + r;
  const double* pl = pe[el];
  double a = 0.0;
#pragma unroll 27
  for (int c = 0; c < 81; c++) a = a + (double)ke[c * 81] * pl[c];
  Y[e * 81 + r] = a;
}

// --------------------------------------------------------------------------
// EBE product on symmetric-packed storage (the fast path)
// --------------------------------------------------------------------------
// K_e keeps its upper triangle U(r,c), r <= c, column by column:
// U(r,c) at c(c+1)/2 + r, 3321 entries padded to SGEO_KPACK = 3328 so each
// element starts 16-byte aligned.  One 128-thread block per element: the
// block stages the packed matrix in LDS with 16-byte loads (each stored
// entry crosses HBM once), then thread r < 81 forms
// y_r = sum_c K(r,c) p_c, c ascending, with K(r,c) = U(min(r,c), max(r,c)).
// The reference's K_e is symmetric only to the last bit, so this path
// agrees with the full-storage product to rounding, not bit for bit.
template <typename KT>
__global__ __launch_bounds__(128) void k_ebe_packed(
    int32_t nelmt, const KT* __restrict__ KP, const int32_t* __restrict__ gdof,
    const double* __restrict__ p, double* __restrict__ Y,
    const PcgScalars* __restrict__ sc, int skip_mode) {
  if (skip_mode == 1 && sc->done) return;
  if (skip_mode == 2 && !sc->u_nonzero) return;
  constexpr int PER = 16 / sizeof(KT);
  constexpr int NV = SGEO_KPACK / PER;
  __shared__ __align__(16) KT u[SGEO_KPACK];
  __shared__ double pe[81];
  const int64_t e = blockIdx.x;
  const int t = threadIdx.x;
  typedef double v2d __attribute__((ext_vector_type(2)));
  typedef float v4f __attribute__((ext_vector_type(4)));
  using V = typename std::conditional<sizeof(KT) == 8, v2d, v4f>::type;
  const V* src = reinterpret_cast<const V*>(KP + e * SGEO_KPACK);
  V* dst = reinterpret_cast<V*>(u);
#pragma unroll 4
  for (int q = t; q < NV; q += 128) dst[q] = __builtin_nontemporal_load(src + q);
  if (t < 81) pe[t] = p[gdof[e * 81 + t]];
  __syncthreads();
  if (t >= 81) return;
  const int r = t;
  const int rbase = r * (r + 1) / 2;
  double a = 0.0;
#pragma unroll 9
  for (int c = 0; c < 81; c++) {
    const int idx = (r <= c) ? (c * (c + 1) / 2 + r) : (rbase + c);
    a = a + (double)u[idx] * pe[c];
  }
  Y[e * 81 + r] = a;
}

// Persistent, pipelined form of k_ebe_packed<double>: a few 128-thread
// blocks per CU walk the elements.  While element e is multiplied from LDS,
// the next element's packed K is in flight into registers (16-byte
// non-temporal loads) and its gathered p into a register, with gdof read
// one element further ahead; both reach LDS after the product, behind the
// block's barrier.  Row r sums c ascending as k_ebe_packed does: the same
// bits.
__global__ __launch_bounds__(128) void k_ebe_packed_pipe(
    int32_t nelmt, const double* __restrict__ KP, const int32_t* __restrict__ gdof,
    const double* __restrict__ p, double* __restrict__ Y,
    const PcgScalars* __restrict__ sc, int skip_mode) {
  if (skip_mode == 1 && sc->done) return;
  if (skip_mode == 2 && !sc->u_nonzero) return;
  typedef double v2d __attribute__((ext_vector_type(2)));
  constexpr int NV = SGEO_KPACK / 2;        // 16-byte pieces per element
  constexpr int PT = (NV + 127) / 128;      // per thread
  __shared__ __align__(16) double u[SGEO_KPACK];
  __shared__ double pe[2][81];
  const int t = threadIdx.x;
  const int64_t stride = gridDim.x;
  int64_t e = blockIdx.x;
  v2d kr[PT];
  int32_t gn = 0;  // gdof of the next element
  if (e < nelmt) {
    const v2d* src = reinterpret_cast<const v2d*>(KP + e * SGEO_KPACK);
#pragma unroll
    for (int i = 0; i < PT; i++) {
      const int q = t + 128 * i;
      if (q < NV) kr[i] = __builtin_nontemporal_load(src + q);
    }
    if (t < 81) pe[0][t] = p[gdof[e * 81 + t]];
    if (t < 81 && e + stride < nelmt) gn = gdof[(e + stride) * 81 + t];
  }
  const int r = t;
  const int rbase = r * (r + 1) / 2;
  int cur = 0;
  for (; e < nelmt; e += stride) {
    v2d* dst = reinterpret_cast<v2d*>(u);
#pragma unroll
    for (int i = 0; i < PT; i++) {
      const int q = t + 128 * i;
      if (q < NV) dst[q] = kr[i];
    }
    __syncthreads();  // K of e and pe[cur] visible
    const int64_t en = e + stride;
    double pn = 0.0;
    if (en < nelmt) {  // in flight during the product
      const v2d* src = reinterpret_cast<const v2d*>(KP + en * SGEO_KPACK);
#pragma unroll
      for (int i = 0; i < PT; i++) {
        const int q = t + 128 * i;
        if (q < NV) kr[i] = __builtin_nontemporal_load(src + q);
      }
      if (t < 81) pn = p[gn];
    }
    int32_t gnn = 0;
    if (t < 81 && en + stride < nelmt) gnn = gdof[(en + stride) * 81 + t];
    if (t < 81) {
      const double* pl = pe[cur];
      double a = 0.0;
#pragma unroll 9
      for (int c = 0; c < 81; c++) {
        const int idx = (r <= c) ? (c * (c + 1) / 2 + r) : (rbase + c);
        a = a + u[idx] * pl[c];
      }
      Y[e * 81 + r] = a;
      pe[cur ^ 1][t] = pn;
    }
    gn = gnn;
    __syncthreads();  // u and pe[cur] free, pe[cur ^ 1] visible
    cur ^= 1;
  }
}

// blocks per CU of the persistent FP64 product (LDS 28 KB per block)
#define EBE_PIPE_BLOCKS_PER_CU 5

// full (reference layout) -> packed upper triangle, zero padding
template <typename KT>
__global__ void k_pack_upper(int32_t nelmt, const double* __restrict__ K,
                             KT* __restrict__ KP) {
  const int64_t n = (int64_t)nelmt * SGEO_KPACK;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = q / SGEO_KPACK;
    const int i = (int)(q - e * SGEO_KPACK);
    KT v = (KT)0;
    if (i < 3321) {
      // column c with c(c+1)/2 <= i < (c+1)(c+2)/2, row r = i - c(c+1)/2
      int c = (int)((sqrt(8.0 * i + 1.0) - 1.0) * 0.5);
      while (c * (c + 1) / 2 > i) c--;
      while ((c + 1) * (c + 2) / 2 <= i) c++;
      const int r = i - c * (c + 1) / 2;
      v = (KT)K[e * 6561 + r + 81 * c];
    }
    KP[q] = v;
  }
}

// plain copy of n doubles (bandwidth calibration)
__global__ __launch_bounds__(256) void k_copy(int64_t n, const double2* __restrict__ a,
                                              double2* __restrict__ b) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    b[q] = a[q];
}

// --------------------------------------------------------------------------
// assembly: out_g = sum over the element slots of g, ascending element order
// (the order of kp(egdof)=kp(egdof)+... in src/solver.f90:54-58)
// --------------------------------------------------------------------------
__device__ __forceinline__ double gather_row(const int32_t* __restrict__ rowptr,
                                             const int32_t* __restrict__ slots,
                                             const double* __restrict__ Y,
                                             int64_t g) {
  double acc = 0.0;
  const int32_t b = rowptr[g], en = rowptr[g + 1];
  for (int32_t s = b; s < en; s++) acc = acc + Y[slots[s]];
  return acc;
}

__global__ __launch_bounds__(SGEO_VEC_THREADS) void k_assemble(
    int64_t n1, const int32_t* __restrict__ rowptr,
    const int32_t* __restrict__ slots, const double* __restrict__ Y,
    double* __restrict__ out, const PcgScalars* __restrict__ sc,
    int skip_mode) {
  if (skip_mode == 2 && !sc->u_nonzero) return;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n1;
       g += (int64_t)gridDim.x * blockDim.x)
    out[g] = (g == 0) ? 0.0 : gather_row(rowptr, slots, Y, g);
}

// out_g = out_g (+|-) its Y terms in element order; out[0] = 0
// (bodyload(egdof)=bodyload(egdof)+bload, src/slope.F90:382;
//  extload(egdof)=extload(egdof)-eld*gam, src/preprocess.f90:136)
__global__ __launch_bounds__(SGEO_VEC_THREADS) void k_assemble_acc(
    int64_t n1, const int32_t* __restrict__ rowptr,
    const int32_t* __restrict__ slots, const double* __restrict__ Y,
    double* __restrict__ out, int subtract,
    const PlasticScalars* __restrict__ ps) {
  if (ps && ps->nl_isconv) return;  // no body load once converged
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n1;
       g += (int64_t)gridDim.x * blockDim.x) {
    if (g == 0) { out[0] = 0.0; continue; }
    double acc = out[g];
    const int32_t b = rowptr[g], en = rowptr[g + 1];
    if (subtract) {
      for (int32_t s = b; s < en; s++) acc = acc - Y[slots[s]];
    } else {
      for (int32_t s = b; s < en; s++) acc = acc + Y[slots[s]];
    }
    out[g] = acc;
  }
}

// --------------------------------------------------------------------------
// PCG vector kernels: partial sums per block
// --------------------------------------------------------------------------
#define VEC_LOOP(g)                                                         \
  const int64_t base = (int64_t)blockIdx.x * SGEO_VEC_THREADS * SGEO_VEC_ITEMS; \
  _Pragma("unroll") for (int k_ = 0; k_ < SGEO_VEC_ITEMS; k_++)            \
    for (int64_t g = base + threadIdx.x + (int64_t)k_ * SGEO_VEC_THREADS;  \
         g < n1; g = n1)

// max|u| at entry (src/solver.f90:39)
__global__ __launch_bounds__(SGEO_VEC_THREADS) void k_pcg_begin(
    int64_t n1, const double* __restrict__ u, double* __restrict__ part) {
  __shared__ double red[SGEO_VEC_THREADS / 64];
  double m = 0.0;
  VEC_LOOP(g) { m = fmax(m, fabs(u[g])); }
  m = block_max<SGEO_VEC_THREADS / 64>(m, red);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// r = f - kp (kp = K u if u != 0, else 0); z = dinv*r; p = z; partial r.z
__global__ __launch_bounds__(SGEO_VEC_THREADS) void k_pcg_init(
    int64_t n1, const double* __restrict__ f, const double* __restrict__ kp,
    const double* __restrict__ dinv, double* __restrict__ r,
    double* __restrict__ z, double* __restrict__ p, double* __restrict__ part,
    const PcgScalars* __restrict__ sc) {
  __shared__ double red[SGEO_VEC_THREADS / 64];
  const bool unz = sc->u_nonzero != 0;
  double s = 0.0;
  VEC_LOOP(g) {
    const double rg = f[g] - (unz ? kp[g] : 0.0);
    const double zg = dinv[g] * rg;
    r[g] = rg;
    z[g] = zg;
    p[g] = zg;
    s = s + rg * zg;
  }
  s = block_sum<SGEO_VEC_THREADS / 64>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// kp = assemble(Y); partial p.kp and max|p|
__global__ __launch_bounds__(SGEO_VEC_THREADS) void k_pcg_assemble_dot(
    int64_t n1, const int32_t* __restrict__ rowptr,
    const int32_t* __restrict__ slots, const double* __restrict__ Y,
    const double* __restrict__ p, double* __restrict__ kp,
    double* __restrict__ part, const PcgScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double red[SGEO_VEC_THREADS / 64];
  double s = 0.0, m = 0.0;
  VEC_LOOP(g) {
    const double v = (g == 0) ? 0.0 : gather_row(rowptr, slots, Y, g);
    kp[g] = v;
    const double pg = p[g];
    s = s + pg * v;
    m = fmax(m, fabs(pg));
  }
  const double bs = block_sum<SGEO_VEC_THREADS / 64>(s, red);
  const double bm = block_max<SGEO_VEC_THREADS / 64>(m, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = bs;
    part[gridDim.x + blockIdx.x] = bm;
  }
}

// u += alpha p; partial max|u|; r -= alpha kp; z = dinv r; partial r.z
__global__ __launch_bounds__(SGEO_VEC_THREADS) void k_pcg_update(
    int64_t n1, const double* __restrict__ p, const double* __restrict__ kp,
    const double* __restrict__ dinv, double* __restrict__ u,
    double* __restrict__ r, double* __restrict__ z, double* __restrict__ part,
    const PcgScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double red[SGEO_VEC_THREADS / 64];
  const double alpha = sc->alpha;
  double m = 0.0, s = 0.0;
  VEC_LOOP(g) {
    const double ug = u[g] + alpha * p[g];
    u[g] = ug;
    m = fmax(m, fabs(ug));
    const double rg = r[g] - alpha * kp[g];
    const double zg = dinv[g] * rg;
    r[g] = rg;
    z[g] = zg;
    s = s + rg * zg;
  }
  const double bm = block_max<SGEO_VEC_THREADS / 64>(m, red);
  const double bs = block_sum<SGEO_VEC_THREADS / 64>(s, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = bm;
    part[gridDim.x + blockIdx.x] = bs;
  }
}

// p = z + beta p
__global__ __launch_bounds__(SGEO_VEC_THREADS) void k_pcg_pupdate(
    int64_t n1, const double* __restrict__ z, double* __restrict__ p,
    const PcgScalars* __restrict__ sc) {
  if (sc->done) return;
  const double beta = sc->beta;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n1;
       g += (int64_t)gridDim.x * blockDim.x)
    p[g] = z[g] + beta * p[g];
}

// --------------------------------------------------------------------------
// finalize: one block combines the nb partials in index order
// --------------------------------------------------------------------------
enum FinWhat {
  FIN_BEGIN = 0,   // part[0..nb): max|u|           -> u_nonzero, reset
  FIN_RZ = 1,      // part[0..nb): r.z              -> rz
  FIN_ALPHA = 2,   // p.kp, max|p|                  -> alpha
  FIN_MAXP = 3,    // max|p| only (p.kp in reference order)
  FIN_UPDATE = 4,  // max|u|, r.z                   -> stop test, beta
  FIN_UPDATE_NOBETA = 5,  // stop test only (r.z in reference order)
};

__global__ __launch_bounds__(256) void k_pcg_finalize(
    int nb, const double* __restrict__ part, PcgScalars* __restrict__ sc,
    int what, double tol, int32_t maxiter) {
  __shared__ double red[4];
  if (what >= FIN_ALPHA && sc->done) return;
  double v0 = 0.0, v1 = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) {
    const double a = part[b];
    if (what == FIN_BEGIN) v0 = fmax(v0, a);
    else if (what == FIN_UPDATE || what == FIN_UPDATE_NOBETA) v0 = fmax(v0, a);
    else if (what != FIN_MAXP) v0 = v0 + a;
    if (what == FIN_ALPHA || what == FIN_MAXP) v1 = fmax(v1, part[nb + b]);
    if (what == FIN_UPDATE) v1 = v1 + part[nb + b];
  }
  const bool v0max = (what == FIN_BEGIN || what == FIN_UPDATE ||
                      what == FIN_UPDATE_NOBETA);
  v0 = v0max ? block_max<4>(v0, red) : block_sum<4>(v0, red);
  const bool v1max = (what == FIN_ALPHA || what == FIN_MAXP);
  v1 = v1max ? block_max<4>(v1, red) : block_sum<4>(v1, red);
  if (threadIdx.x != 0) return;
  switch (what) {
    case FIN_BEGIN:
      sc->u_nonzero = v0 > 0.0 ? 1 : 0;
      sc->iter = 0;
      sc->done = 0;
      sc->conv_iter = 0;
      sc->tol = tol;
      sc->maxiter = maxiter;
      break;
    case FIN_RZ:
      sc->rz = v0;
      break;
    case FIN_ALPHA:
      sc->pkp = v0;
      sc->maxp = v1;
      sc->alpha = sc->rz / v0;
      break;
    case FIN_MAXP:
      sc->maxp = v1;
      break;
    case FIN_UPDATE:
    case FIN_UPDATE_NOBETA: {
      const int it = sc->iter + 1;
      sc->iter = it;
      sc->maxu = v0;
      // src/solver.f90:65: stop after the u update, before the r update
      if (fabs(sc->alpha) * sc->maxp / v0 <= sc->tol) {
        sc->done = 1;
        sc->conv_iter = it;
      } else if (it >= sc->maxiter) {
        sc->done = 2;
        sc->conv_iter = sc->maxiter + 1;  // DO-loop variable after exhaustion
      } else if (what == FIN_UPDATE) {
        sc->beta = v1 / sc->rz;
        sc->rz = v1;
      }
      break;
    }
  }
}

// --------------------------------------------------------------------------
// reference-order dot products: s = ((0 + a0 b0) + a1 b1) + ..., exactly the
// sequential dot_product of src/solver.f90:61-62,72.  One block: waves 1-3
// stage rounded products of the next chunk in LDS while lane 0 of wave 0
// adds the current chunk in index order.  O(n) latency-bound; meant for
// bit-for-bit parity runs, the tree reductions above are the fast path.
// --------------------------------------------------------------------------
#define SEQ_CHUNK 1536
__device__ double seq_dot(int64_t n1, const double* __restrict__ a,
                          const double* __restrict__ b, double* buf) {
  const int t = threadIdx.x;
  const int64_t nch = (n1 + SEQ_CHUNK - 1) / SEQ_CHUNK;
  double s = 0.0;
  for (int j = t; j < SEQ_CHUNK; j += blockDim.x)
    buf[j] = (j < n1) ? a[j] * b[j] : 0.0;
  __syncthreads();
  for (int64_t c = 0; c < nch; c++) {
    const double* cur = buf + (c & 1) * SEQ_CHUNK;
    if (t < 64) {
      if (t == 0) {
        const int64_t cnt = (c == nch - 1) ? n1 - c * SEQ_CHUNK : SEQ_CHUNK;
        for (int64_t j = 0; j < cnt; j++) s = s + cur[j];
      }
    } else if (c + 1 < nch) {
      double* nxt = buf + ((c + 1) & 1) * SEQ_CHUNK;
      const int64_t base = (c + 1) * SEQ_CHUNK;
      for (int j = t - 64; j < SEQ_CHUNK; j += blockDim.x - 64) {
        const int64_t g = base + j;
        nxt[j] = (g < n1) ? a[g] * b[g] : 0.0;
      }
    }
    __syncthreads();
  }
  return s;  // valid in thread 0
}

// which: 0 initial rz; 1 alpha = rz / p.kp; 2 beta from r.z
__global__ __launch_bounds__(256) void k_pcg_seqdot(
    int64_t n1, const double* __restrict__ a, const double* __restrict__ b,
    PcgScalars* __restrict__ sc, int which) {
  __shared__ double buf[2 * SEQ_CHUNK];
  if (which != 0 && sc->done) return;
  const double s = seq_dot(n1, a, b, buf);
  if (threadIdx.x == 0) {
    if (which == 0) {
      sc->rz = s;
    } else if (which == 1) {
      sc->pkp = s;
      sc->alpha = sc->rz / s;
    } else {
      sc->beta = s / sc->rz;
      sc->rz = s;
    }
  }
}

// --------------------------------------------------------------------------
// plastic-iteration vector kernels (src/slope.F90:288-309)
// --------------------------------------------------------------------------
// load = extload + bodyload; load(0) = 0
__global__ void k_load(int64_t n1, const double* __restrict__ ext,
                       const double* __restrict__ body, double* __restrict__ load) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n1;
       g += (int64_t)gridDim.x * blockDim.x)
    load[g] = (g == 0) ? 0.0 : ext[g] + body[g];
}

// x(0)=0; partial max|x-oldx| and max|x|; oldx = x
__global__ __launch_bounds__(SGEO_VEC_THREADS) void k_uerr(
    int64_t n1, double* __restrict__ x, double* __restrict__ oldx,
    double* __restrict__ part) {
  __shared__ double red[SGEO_VEC_THREADS / 64];
  double md = 0.0, mx = 0.0;
  VEC_LOOP(g) {
    double xg = x[g];
    if (g == 0) { xg = 0.0; x[0] = 0.0; }
    md = fmax(md, fabs(xg - oldx[g]));
    mx = fmax(mx, fabs(xg));
    oldx[g] = xg;
  }
  const double bd = block_max<SGEO_VEC_THREADS / 64>(md, red);
  const double bx = block_max<SGEO_VEC_THREADS / 64>(mx, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = bd;
    part[gridDim.x + blockIdx.x] = bx;
  }
}

// uerr = max|x-oldx| / max|x|; nl_isconv = uerr <= nl_tol; fmax = 0
__global__ __launch_bounds__(256) void k_uerr_finalize(
    int nb, const double* __restrict__ part, PlasticScalars* __restrict__ ps,
    double nl_tol) {
  __shared__ double red[4];
  double vd = 0.0, vx = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) {
    vd = fmax(vd, part[b]);
    vx = fmax(vx, part[nb + b]);
  }
  vd = block_max<4>(vd, red);
  vx = block_max<4>(vx, red);
  if (threadIdx.x == 0) {
    const double uerr = vd / vx;
    ps->uerr = uerr;
    ps->xmax = vx;
    ps->nl_isconv = uerr <= nl_tol ? 1 : 0;
    ps->fmax_bits = 0ull;  // fmax = zero (src/slope.F90:286)
    ps->n_plastic = 0;
  }
}

// --------------------------------------------------------------------------
// inverse map (equation -> element slots) built on the device: the same
// CSR the host counting sort of sgeo_mesh_create builds (slots of every
// equation g >= 1 in ascending e*81+i order; the index-0 sink stays empty,
// no assembly reads it).  Every kernel checks the equation range itself.
// --------------------------------------------------------------------------
__global__ void k_fill_i32(int64_t n, int32_t* __restrict__ a, int32_t v) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    a[q] = v;
}

// cnt[g + 1] = slots of equation g (1 <= g <= neq); *bad = 1 if an entry
// lies outside 0..neq (such entries are skipped)
__global__ void k_csr_count(int64_t ns, int32_t neq, const int32_t* __restrict__ gdof,
                            int32_t* __restrict__ cnt, int32_t* __restrict__ bad) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < ns;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int32_t g = gdof[q];
    if (g < 0 || g > neq) {
      *bad = 1;
      continue;
    }
    if (g > 0) atomicAdd(&cnt[g + 1], 1);
  }
}

// in-place inclusive scan of a[0..n) by one block of 1024 threads, 4096
// entries per round (4 per thread), the running total carried between rounds
__global__ __launch_bounds__(1024) void k_scan_i32(int64_t n, int32_t* __restrict__ a) {
  __shared__ int32_t wsum[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int32_t carry = 0;
  for (int64_t base = 0; base < n; base += 4096) {
    const int64_t i0 = base + (int64_t)t * 4;
    int32_t v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      s += (i0 + k < n) ? a[i0 + k] : 0;
      v[k] = s;
    }
    int32_t x = s;  // inclusive scan of the thread totals across the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
      int32_t ws = lane < 16 ? wsum[lane] : 0;
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        const int32_t y = __shfl_up(ws, off, 64);
        if (lane >= off) ws += y;
      }
      if (lane < 16) wsum[lane] = ws;
    }
    __syncthreads();
    const int32_t excl = carry + (x - s) + (w > 0 ? wsum[w - 1] : 0);
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (i0 + k < n) a[i0 + k] = v[k] + excl;
    carry += wsum[15];
    __syncthreads();  // wsum is rewritten by the next round
  }
}

// cursor[g] = rowptr[g]; slots[cursor[g]++] = q for every entry of g >= 1
__global__ void k_csr_fill(int64_t ns, int32_t neq, const int32_t* __restrict__ gdof,
                           int32_t* __restrict__ cursor, int32_t* __restrict__ slots) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < ns;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int32_t g = gdof[q];
    if (g < 1 || g > neq) continue;
    slots[atomicAdd(&cursor[g], 1)] = (int32_t)q;
  }
}

// each equation's slots in ascending order (the fill's atomics leave them
// in arrival order); rows hold at most a node's element count
__global__ void k_csr_sort(int64_t n1, const int32_t* __restrict__ rowptr,
                           int32_t* __restrict__ slots) {
  for (int64_t g = 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n1;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int32_t b = rowptr[g], e = rowptr[g + 1];
    for (int32_t i = b + 1; i < e; i++) {
      const int32_t v = slots[i];
      int32_t j = i - 1;
      while (j >= b && slots[j] > v) {
        slots[j + 1] = slots[j];
        j--;
      }
      slots[j + 1] = v;
    }
  }
}

__global__ void k_copy_i32(int64_t n, const int32_t* __restrict__ a, int32_t* __restrict__ b) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    b[q] = a[q];
}

// *a = v, queued on the stream (a kernel, not a memset node)
void launch_set_i32(hipStream_t s, int32_t* a, int32_t v) {
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, (int64_t)1, a, v);
}

// rowptr (n1 + 1 entries) and slots of gdof (ns entries) into caller memory;
// cursor (n1 entries) is a temporary; *bad (zeroed here) reports entries
// outside 0..neq
void launch_csr_build(hipStream_t s, int64_t ns, int32_t neq, const int32_t* gdof,
                      int32_t* rowptr, int32_t* slots, int32_t* cursor, int32_t* bad) {
  const int64_t n1 = (int64_t)neq + 1;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_stride_blocks(n1 + 1)), dim3(SGEO_VEC_THREADS),
                     0, s, n1 + 1, rowptr, 0);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, (int64_t)1, bad, 0);
  hipLaunchKernelGGL(k_csr_count, dim3(grid_stride_blocks(ns)), dim3(SGEO_VEC_THREADS), 0,
                     s, ns, neq, gdof, rowptr, bad);
  hipLaunchKernelGGL(k_scan_i32, dim3(1), dim3(1024), 0, s, n1 + 1, rowptr);
  hipLaunchKernelGGL(k_copy_i32, dim3(grid_stride_blocks(n1)), dim3(SGEO_VEC_THREADS), 0,
                     s, n1, rowptr, cursor);
  hipLaunchKernelGGL(k_csr_fill, dim3(grid_stride_blocks(ns)), dim3(SGEO_VEC_THREADS), 0,
                     s, ns, neq, gdof, cursor, slots);
  hipLaunchKernelGGL(k_csr_sort, dim3(grid_stride_blocks(n1)), dim3(SGEO_VEC_THREADS), 0,
                     s, n1, rowptr, slots);
}

// --------------------------------------------------------------------------
// launch wrappers
// --------------------------------------------------------------------------
void launch_ebe_elem_f64(hipStream_t s, int32_t nelmt, const double* K,
                         const int32_t* gdof, const double* p, double* Y,
                         const PcgScalars* sc, int skip_mode) {
  if (nelmt <= 0) return;
  hipLaunchKernelGGL(k_ebe_elem<double>, dim3(nblk(nelmt, 3)), dim3(256), 0, s,
                     nelmt, K, gdof, p, Y, sc, skip_mode);
}

void launch_ebe_elem_f32k(hipStream_t s, int32_t nelmt, const float* K,
                          const int32_t* gdof, const double* p, double* Y,
                          const PcgScalars* sc, int skip_mode) {
  if (nelmt <= 0) return;
  hipLaunchKernelGGL(k_ebe_elem<float>, dim3(nblk(nelmt, 3)), dim3(256), 0, s,
                     nelmt, K, gdof, p, Y, sc, skip_mode);
}

void launch_ebe_packed_f64(hipStream_t s, int32_t nelmt, int n_cu, const double* KP,
                           const int32_t* gdof, const double* p, double* Y,
                           const PcgScalars* sc, int skip_mode) {
  if (nelmt <= 0) return;
  int64_t g = (int64_t)n_cu * EBE_PIPE_BLOCKS_PER_CU;
  if (g > nelmt) g = nelmt;
  hipLaunchKernelGGL(k_ebe_packed_pipe, dim3((unsigned)g), dim3(128), 0, s, nelmt, KP,
                     gdof, p, Y, sc, skip_mode);
}

void launch_ebe_packed_f32(hipStream_t s, int32_t nelmt, const float* KP,
                           const int32_t* gdof, const double* p, double* Y,
                           const PcgScalars* sc, int skip_mode) {
  if (nelmt <= 0) return;
  hipLaunchKernelGGL(k_ebe_packed<float>, dim3(nelmt), dim3(128), 0, s, nelmt,
                     KP, gdof, p, Y, sc, skip_mode);
}

void launch_pack_upper_f64(hipStream_t s, int32_t nelmt, const double* K,
                           double* KP) {
  hipLaunchKernelGGL(k_pack_upper<double>,
                     dim3(grid_stride_blocks((int64_t)nelmt * SGEO_KPACK)),
                     dim3(SGEO_VEC_THREADS), 0, s, nelmt, K, KP);
}

void launch_pack_upper_f32(hipStream_t s, int32_t nelmt, const double* K,
                           float* KP) {
  hipLaunchKernelGGL(k_pack_upper<float>,
                     dim3(grid_stride_blocks((int64_t)nelmt * SGEO_KPACK)),
                     dim3(SGEO_VEC_THREADS), 0, s, nelmt, K, KP);
}

void launch_copy(hipStream_t s, int64_t n_double2, const double* a, double* b) {
  hipLaunchKernelGGL(k_copy, dim3(grid_stride_blocks(n_double2)),
                     dim3(256), 0, s, n_double2, (const double2*)a, (double2*)b);
}

void launch_assemble(hipStream_t s, int64_t n1, const int32_t* rowptr,
                     const int32_t* slots, const double* Y, double* out,
                     const PcgScalars* sc, int skip_mode) {
  hipLaunchKernelGGL(k_assemble, dim3(grid_stride_blocks(n1)),
                     dim3(SGEO_VEC_THREADS), 0, s, n1, rowptr, slots, Y, out,
                     sc, skip_mode);
}

void launch_assemble_acc(hipStream_t s, int64_t n1, const int32_t* rowptr,
                         const int32_t* slots, const double* Y, double* out,
                         int subtract, const PlasticScalars* ps) {
  hipLaunchKernelGGL(k_assemble_acc, dim3(grid_stride_blocks(n1)),
                     dim3(SGEO_VEC_THREADS), 0, s, n1, rowptr, slots, Y, out,
                     subtract, ps);
}

void launch_pcg_begin(hipStream_t s, int64_t n1, const double* u, double* part,
                      PcgScalars* sc, double tol, int32_t maxiter) {
  const unsigned nb = vec_reduce_blocks(n1);
  hipLaunchKernelGGL(k_pcg_begin, dim3(nb), dim3(SGEO_VEC_THREADS), 0, s, n1, u,
                     part);
  hipLaunchKernelGGL(k_pcg_finalize, dim3(1), dim3(256), 0, s, (int)nb, part, sc,
                     (int)FIN_BEGIN, tol, maxiter);
}

void launch_pcg_init(hipStream_t s, int64_t n1, const double* f,
                     const double* kp, const double* dinv, double* r, double* z,
                     double* p, double* part, PcgScalars* sc, int seq) {
  const unsigned nb = vec_reduce_blocks(n1);
  hipLaunchKernelGGL(k_pcg_init, dim3(nb), dim3(SGEO_VEC_THREADS), 0, s, n1, f,
                     kp, dinv, r, z, p, part, sc);
  if (seq)
    hipLaunchKernelGGL(k_pcg_seqdot, dim3(1), dim3(256), 0, s, n1, r, z, sc, 0);
  else
    hipLaunchKernelGGL(k_pcg_finalize, dim3(1), dim3(256), 0, s, (int)nb, part,
                       sc, (int)FIN_RZ, 0.0, 0);
}

void launch_pcg_assemble_dot(hipStream_t s, int64_t n1, const int32_t* rowptr,
                             const int32_t* slots, const double* Y,
                             const double* p, double* kp, double* part,
                             PcgScalars* sc, int seq) {
  const unsigned nb = vec_reduce_blocks(n1);
  hipLaunchKernelGGL(k_pcg_assemble_dot, dim3(nb), dim3(SGEO_VEC_THREADS), 0, s,
                     n1, rowptr, slots, Y, p, kp, part, sc);
  hipLaunchKernelGGL(k_pcg_finalize, dim3(1), dim3(256), 0, s, (int)nb, part, sc,
                     (int)(seq ? FIN_MAXP : FIN_ALPHA), 0.0, 0);
  if (seq)
    hipLaunchKernelGGL(k_pcg_seqdot, dim3(1), dim3(256), 0, s, n1, p, kp, sc, 1);
}

void launch_pcg_update(hipStream_t s, int64_t n1, const double* p,
                       const double* kp, const double* dinv, double* u,
                       double* r, double* z, double* part, PcgScalars* sc,
                       int seq) {
  const unsigned nb = vec_reduce_blocks(n1);
  hipLaunchKernelGGL(k_pcg_update, dim3(nb), dim3(SGEO_VEC_THREADS), 0, s, n1, p,
                     kp, dinv, u, r, z, part, sc);
  hipLaunchKernelGGL(k_pcg_finalize, dim3(1), dim3(256), 0, s, (int)nb, part, sc,
                     (int)(seq ? FIN_UPDATE_NOBETA : FIN_UPDATE), 0.0, 0);
  if (seq)
    hipLaunchKernelGGL(k_pcg_seqdot, dim3(1), dim3(256), 0, s, n1, r, z, sc, 2);
}

void launch_pcg_pupdate(hipStream_t s, int64_t n1, const double* z, double* p,
                        const PcgScalars* sc) {
  hipLaunchKernelGGL(k_pcg_pupdate, dim3(grid_stride_blocks(n1)),
                     dim3(SGEO_VEC_THREADS), 0, s, n1, z, p, sc);
}

void launch_load(hipStream_t s, int64_t n1, const double* ext,
                 const double* body, double* load) {
  hipLaunchKernelGGL(k_load, dim3(grid_stride_blocks(n1)),
                     dim3(SGEO_VEC_THREADS), 0, s, n1, ext, body, load);
}

void launch_uerr(hipStream_t s, int64_t n1, double* x, double* oldx,
                 double* part, PlasticScalars* ps, double nl_tol) {
  const unsigned nb = vec_reduce_blocks(n1);
  hipLaunchKernelGGL(k_uerr, dim3(nb), dim3(SGEO_VEC_THREADS), 0, s, n1, x, oldx,
                     part);
  hipLaunchKernelGGL(k_uerr_finalize, dim3(1), dim3(256), 0, s, (int)nb, part, ps,
                     nl_tol);
}

}  // namespace sgeo
