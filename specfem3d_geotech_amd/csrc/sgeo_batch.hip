// sgeo_batch.hip -- multi-trial (multi-RHS) kernels of the mixed-precision
// EBE-PCG: B strength-reduction trials share one mesh and one elastic
// stiffness (the viscoplastic method never changes K, src/slope.F90:288), so
// every CG iteration streams K once for all B columns.
//
// Vectors are interleaved [n1][B] (column fastest): an element gathers 81
// contiguous B-vectors and every update is unit-stride.  The inner CG
// vectors (r, p) and the element product are FP32 (the product on MFMA; K p
// is summed from the element products inside the update, never stored);
// dot products, the solution x, the FP64 true residual f - K x and all
// plastic state are FP64.  Every reduction combines per-chunk (or
// per-element) partials in groups, then per column, in a fixed order, so a
// column's arithmetic does not depend on B or on the other columns: a trial
// gives the same bits alone and in any batch.
#pragma clang fp contract(off)

#include "sgeo_device.h"
#include "sgeo_internal.h"

namespace sgeo {

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bool wanted(const int32_t* cst, int j, int mask) {
  return (mask >> cst[j]) & 1;
}

// --------------------------------------------------------------------------
// FP32 batched element product on MFMA:
//   Y[e][r][j] = sum_c K_e(r,c) P[g(e,c)][j]
// with v_mfma_f32_16x16x4f32 (A 16x4, B 4x16, D 16x16): rows 0..79 in 5
// tiles of 16, columns in tiles of 16 (B < 16: one partial tile), k in 21
// steps of 4 (81 -> 84).  Row 80 is a dot product on the vector ALUs: lane
// (j, kq) sums K(80, 4s+kq) P(4s+kq, j) over the k-steps s -- the B operand
// it already holds -- and the four kq partials are added across lanes in a
// fixed order.  A sixth MFMA row tile would spend 15 of its 16 rows on
// zero padding.  One element per block of 4 waves; wave w takes column
// tile n = w % NT and row tiles m = w / NT (+ 4 / NT ...); the wave of each
// column tile with the fewest row tiles also takes row 80.
// The packed upper triangle of K_e is unpacked (with its mirror) into LDS
// row by row, Kf[r * BMV_RLD + slot(c)] with slot(c) = (c % 4) * 24 + c / 4:
// lane (i, kq) of k-steps 4m .. 4m+3 needs K(r, 4s + kq) for s = 4m .. 4m+3,
// which sit at four consecutive slots, so one 16-byte read serves four
// MFMAs; zero padded for k = 81..83.  BMV_RLD = 100 floats puts the 16 rows
// of a read on disjoint groups of four banks and keeps every read 16-byte
// aligned.
// A column's result depends only on its own P column, so trials are
// independent of the batch; columns not in `want_mask` are not written.
//
// The epilogue also forms the element's share of p.Kp,
//   edot[e][j] = sum_r P(g(e,r), j) Y[e][r][j]   (FP64 sums of exact products),
// since p.Kp = sum_g p_g sum_{e,r: g(e,r)=g} Y[e][r] = sum_e p_e . Y_e.  The
// accumulator of row tile mt holds Y(16 mt + 4 kq + v, j) in lane (kq, j),
// register v, while the B operand of k-step 4 mt + i holds P(16 mt + 4 i +
// kq, j) in lane (kq, j).  A 4x4 transpose of (lane group, register) -- two
// v_permlane32_swap and two v_permlane16_swap -- puts Y(16 mt + 4 i + kq, j)
// in register i of lane (kq, j), beside its P.  The share is summed in one
// order for every column tiling: per lane, row tiles 0..4 (each the four
// products i = 0..3 in order), then row 80 in lane group 0, then the four
// lane groups (xor 16, xor 32; IEEE addition commutes, so all four lanes
// hold the same sum).  When a column tile's row tiles are spread over
// several waves (B <= 32), the per-tile sums meet in LDS and one wave adds
// them in that order.  Per-element shares keep the reduction independent of
// which block computed an element.
// --------------------------------------------------------------------------
#define BMV_RLD 100
#define BMV_KC 84
__device__ __forceinline__ constexpr int bmv_slot(int c) { return (c & 3) * 24 + (c >> 2); }

// (lane group, register) transpose of a 4x4 block of 16-lane groups:
// on return a[i] in lane group g holds what a[g] held in lane group i
__device__ __forceinline__ void lanegroup_transpose4(float (&a)[4]) {
  uint32_t u0 = __float_as_uint(a[0]), u1 = __float_as_uint(a[1]);
  uint32_t u2 = __float_as_uint(a[2]), u3 = __float_as_uint(a[3]);
  // lanes 32..63 of the first operand <-> lanes 0..31 of the second
  auto s02 = __builtin_amdgcn_permlane32_swap(u0, u2, false, false);
  auto s13 = __builtin_amdgcn_permlane32_swap(u1, u3, false, false);
  // odd 16-lane rows of the first operand <-> even rows of the second
  auto t01 = __builtin_amdgcn_permlane16_swap(s02[0], s13[0], false, false);
  auto t23 = __builtin_amdgcn_permlane16_swap(s02[1], s13[1], false, false);
  a[0] = __uint_as_float(t01[0]);
  a[1] = __uint_as_float(t01[1]);
  a[2] = __uint_as_float(t23[0]);
  a[3] = __uint_as_float(t23[1]);
}

__device__ __forceinline__ int tri_col(int q) {  // c with c(c+1)/2 <= q < (c+1)(c+2)/2
  int c = (int)((sqrtf(8.0f * (float)q + 1.0f) - 1.0f) * 0.5f);
  if ((c + 1) * (c + 2) / 2 <= q) c++;
  if (c * (c + 1) / 2 > q) c--;
  return c;
}

// Persistent and software-pipelined: the grid holds a few blocks per CU.
// The elements are taken in the order eorder[] (spatially compact runs,
// built with the engine).  With nxcd = 8, each XCD (blocks are dispatched
// to the 8 XCDs round-robin) walks its own contiguous eighth of that order,
// so elements in flight together share nodes and their gathered P rows hit
// in that XCD's L2; nxcd = 1 is a plain grid-stride walk.  While element e is multiplied, K
// and gdof of the block's next element are already in registers, and the
// next element's P rows are gathered before its K is unpacked.  The grid
// size and nxcd change which block computes an element, never how: the
// launch geometry does not change any result bit.
#define BMV_XCD 8

template <int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
void k_bmv_mfma(int32_t nelmt, int B, int nxcd,
                                                  const float* __restrict__ KP,
                                                  const int32_t* __restrict__ gdof,
                                                  const int32_t* __restrict__ eorder,
                                                  const float* __restrict__ P,
                                                  int64_t nv,
                                                  float* __restrict__ Y,
                                                  double* __restrict__ edot,
                                                  const int32_t* __restrict__ cst,
                                                  const BCol* __restrict__ bc,
                                                  int want_mask) {
  constexpr int NS = 4 / NT;             // waves per column tile
  constexpr int NMT = 5;                 // MFMA row tiles: rows 0..79
  constexpr int NM = (NMT + NS - 1) / NS;  // row tiles per wave
  constexpr int KPT = (3321 + 255) / 256;  // packed entries per thread
  __shared__ __align__(16) float Kf[81 * BMV_RLD];
  __shared__ int32_t sg[2][BMV_KC];
  // p.Kp shares of the row tiles (and row 80) of each column tile, when they
  // come from several waves
  __shared__ double xs[NS > 1 ? NT : 1][NS > 1 ? NMT + 1 : 1][64];
  const int t = threadIdx.x;
  // range of the element walk eorder[] of this block's XCD, and the block's
  // stride within it; wi is the walk position, e = eorder[wi] the element
  const int xcd = blockIdx.x % nxcd;
  const int nbx = gridDim.x / nxcd;  // grid is a multiple of nxcd
  const int64_t xe0 = (int64_t)nelmt * xcd / nxcd;
  const int64_t xe1 = (int64_t)nelmt * (xcd + 1) / nxcd;
  int64_t wi = xe0 + blockIdx.x / nxcd;
  int64_t e = wi < xe1 ? eorder[wi] : 0;
  // column tile of this wave; columns outside the mask are neither
  // gathered nor multiplied nor stored, but the wave keeps to the block's
  // barriers
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int l = t & 63;
  const int n = w % NT, mpart = NS == 1 ? 0 : w / NT;
  const int j = n * 16 + (l & 15);
  const bool jv = j < B && wanted(cst, j, want_mask);
  const bool wave_on = __any(jv);
  const int jl = j < B ? j : B - 1;
  // lanes of columns not in the mask (and past B) gather nothing: their B
  // operand is 0 and their results are not stored.  Their own search
  // direction can sit in any of the ring's buffers, so a gather there would
  // fetch lines no wanted column uses.
  // the column's search direction of the iteration being computed
  // (pmask + 1) n1 B < 2^32, checked when the engine is made
  const uint32_t poff = (uint32_t)((bc[jl].it + 1) & bc[jl].pmask) * (uint32_t)nv;
  const int kq = l >> 4;
  const bool row80 = mpart == NS - 1;  // wave-uniform
  // zero padding, once: k columns 81..83 (rows 0..80)
  for (int q = t; q < 3 * 81; q += 256) Kf[(q % 81) * BMV_RLD + bmv_slot(81 + q / 81)] = 0.0f;
  if (t >= 81 && t < BMV_KC) { sg[0][t] = 0; sg[1][t] = 0; }
  // LDS addresses of this thread's packed entries q = t + 256 i (fixed for
  // the kernel): K(r, c) at r*RLD + slot(c) and its mirror at c*RLD +
  // slot(r), both below 2^16, packed in one word
  uint32_t kaddr[KPT];
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const int q = t + 256 * i;
    const int c = q < 3321 ? tri_col(q) : 0;
    const int r = q < 3321 ? q - c * (c + 1) / 2 : 0;
    kaddr[i] = (uint32_t)(r * BMV_RLD + bmv_slot(c)) |
               ((uint32_t)(c * BMV_RLD + bmv_slot(r)) << 16);
  }
  // prologue: K and gdof of the first element
  float kreg[KPT];
  int32_t greg = 0;
  if (wi < xe1) {
    const float* src = KP + e * SGEO_KPACK;
#pragma unroll
    for (int i = 0; i < KPT; i++) {
      const int q = t + 256 * i;
      kreg[i] = q < 3321 ? src[q] : 0.0f;
    }
    if (t < 81) greg = gdof[e * 81 + t];
  }
  if (t < 81) sg[0][t] = greg;
  __syncthreads();
  float bv[21];
  if (wave_on && wi < xe1) {
#pragma unroll
    for (int s = 0; s < 21; s++)
      bv[s] = jv ? P[(uint32_t)(sg[0][4 * s + kq] * B + jl) + poff] : 0.0f;
  }
  int cur = 0;
  for (; wi < xe1; wi += nbx) {
    // unpack the packed upper triangle (and its mirror) of element e
#pragma unroll
    for (int i = 0; i < KPT; i++) {
      if (t + 256 * i < 3321) {
        Kf[kaddr[i] & 0xffffu] = kreg[i];
        Kf[kaddr[i] >> 16] = kreg[i];
      }
    }
    __syncthreads();
    const int64_t wn = wi + nbx;
    const int64_t en = wn < xe1 ? eorder[wn] : 0;
    if (wn < xe1) {  // next element's K and gdof, in flight during the MFMAs
      const float* src = KP + en * SGEO_KPACK;
#pragma unroll
      for (int i = 0; i < KPT; i++) {
        const int q = t + 256 * i;
        kreg[i] = q < 3321 ? src[q] : 0.0f;
      }
      if (t < 81) greg = gdof[en * 81 + t];
    }
    if (wave_on) {
      f32x4 acc[NM];
#pragma unroll
      for (int i = 0; i < NM; i++) acc[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      float r80 = 0.0f;
      // k-steps s = 0..19 four at a time (one 16-byte read per row tile),
      // then s = 20; every accumulator sums its k-steps in order 0..20
      const float* kr = Kf + (l & 15) * BMV_RLD + kq * 24;  // row (l & 15), group kq
#pragma unroll
      for (int m = 0; m < 5; m++) {
        f32x4 a[NM];
        f32x4 a80;
#pragma unroll
        for (int i = 0; i < NM; i++) {
          const int mt = mpart + i * NS;  // wave-uniform
          if (NM * NS > NMT && mt >= NMT) break;
          a[i] = *reinterpret_cast<const f32x4*>(kr + mt * 16 * BMV_RLD + 4 * m);
        }
        if (row80)
          a80 = *reinterpret_cast<const f32x4*>(Kf + 80 * BMV_RLD + kq * 24 + 4 * m);
#pragma unroll
        for (int ss = 0; ss < 4; ss++) {
          const int s = 4 * m + ss;
#pragma unroll
          for (int i = 0; i < NM; i++) {
            const int mt = mpart + i * NS;
            if (NM * NS > NMT && mt >= NMT) break;
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][ss], bv[s], acc[i], 0, 0, 0);
          }
          if (row80) r80 = __builtin_fmaf(a80[ss], bv[s], r80);
        }
      }
#pragma unroll
      for (int i = 0; i < NM; i++) {
        const int mt = mpart + i * NS;
        if (NM * NS > NMT && mt >= NMT) break;
        acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(kr[mt * 16 * BMV_RLD + 20], bv[20],
                                                      acc[i], 0, 0, 0);
      }
      if (row80) r80 = __builtin_fmaf(Kf[80 * BMV_RLD + kq * 24 + 20], bv[20], r80);
      if (row80) {  // the four k-groups of column j: lanes j, j+16, j+32, j+48
        r80 = r80 + __shfl_xor(r80, 16);
        r80 = r80 + __shfl_xor(r80, 32);
      }
      if (jv) {
#pragma unroll
        for (int i = 0; i < NM; i++) {
          const int mt = mpart + i * NS;
          if (NM * NS > NMT && mt >= NMT) break;
#pragma unroll
          for (int v = 0; v < 4; v++)
            Y[(e * 81 + mt * 16 + 4 * kq + v) * B + j] = acc[i][v];
        }
        if (row80 && kq == 0) Y[(e * 81 + 80) * B + j] = r80;
      }
      // the element's share of p.Kp, per lane: row tiles in order, then row 80
      double dsum = 0.0;
#pragma unroll
      for (int i = 0; i < NM; i++) {
#pragma unroll
        for (int mc = 0; mc < NMT; mc++) {
          if (mc != mpart + i * NS) continue;  // wave-uniform; keeps bv indices static
          float a[4] = {acc[i][0], acc[i][1], acc[i][2], acc[i][3]};
          lanegroup_transpose4(a);
          double u = (double)a[0] * (double)bv[4 * mc];
          u = u + (double)a[1] * (double)bv[4 * mc + 1];
          u = u + (double)a[2] * (double)bv[4 * mc + 2];
          u = u + (double)a[3] * (double)bv[4 * mc + 3];
          if (NS == 1) dsum = dsum + u;
          else xs[n][mc][l] = u;
        }
      }
      const double d80 = (row80 && kq == 0) ? (double)r80 * (double)bv[20] : 0.0;
      if (NS == 1) {
        dsum = dsum + d80;
        dsum = dsum + __shfl_xor(dsum, 16);
        dsum = dsum + __shfl_xor(dsum, 32);
        if (jv && kq == 0) edot[e * B + j] = dsum;
      } else if (row80) {
        xs[n][NMT][l] = d80;
      }
    }
    if (t < 81) sg[cur ^ 1][t] = greg;
    __syncthreads();  // Kf free for the next element, its gdof visible
    cur ^= 1;
    if (NS > 1 && wave_on && mpart == 0) {  // the shares of the column tile's waves
      double dsum = 0.0;
#pragma unroll
      for (int mc = 0; mc <= NMT; mc++) dsum = dsum + xs[n][mc][l];
      dsum = dsum + __shfl_xor(dsum, 16);
      dsum = dsum + __shfl_xor(dsum, 32);
      if (jv && kq == 0) edot[e * B + j] = dsum;
    }
    if (wave_on && wn < xe1) {
#pragma unroll
      for (int s = 0; s < 21; s++)
        bv[s] = jv ? P[(uint32_t)(sg[cur][4 * s + kq] * B + jl) + poff] : 0.0f;
    }
    e = en;
  }
}

// --------------------------------------------------------------------------
// FP64 batched element product for the true residual f - K x, on the FP64
// MFMA (v_mfma_f64_16x16x4f64), for the columns of the residual list only:
//   Y64[e][r][q] = sum_c K_e(r,c) X[g(e,c)][rlist[q]],  q < nres = rlist[64],
// with the list's columns packed into tiles of 16 (q = 16 n + lane % 16), so
// the work scales with the columns that need a residual, not with B.  The
// packed upper triangle of K_e (FP64) is unpacked into LDS as
// Kf64[c * B64_LDK + r]; rows 0..79 are 5 row tiles of 16, k runs in 21
// steps of 4 (81 -> 84, zero padded), row 80 is a dot product on the vector
// ALUs as in the FP32 product.  The list length is read on the device, so
// one captured launch serves any list.  Operand layouts: A (16x4) lane l
// holds A(l % 16, l / 16); B (4x16) lane l holds B(l / 16, l % 16); D (16x16)
// lane l, register v holds D(4 v + l / 16, l % 16) -- rows interleaved,
// unlike the FP32 instruction's 4 (l / 16) + v.  A column's sums do not
// depend on its place in the list.
// Persistent and pipelined like the FP32 product: K and gdof of the block's
// next element are loaded during the MFMAs.
// --------------------------------------------------------------------------
#define B64_LDK 88
typedef double f64x4 __attribute__((ext_vector_type(4)));

// Wave w multiplies list tile w (list places 16 w .. 16 w + 15) over all
// row tiles; waves past the list idle.
__device__ __forceinline__ void bmv64_element(
    int64_t e, int nres, int B, const double* __restrict__ X, double* __restrict__ Y,
    const int32_t* __restrict__ rlist, const double* Kf, const int32_t* sgc, int w, int l) {
  constexpr int NMT = 5;
  if (w * 16 >= nres) return;  // wave-uniform: no column in this tile
  const int q = w * 16 + (l & 15);  // list place of this lane's column
  const bool qv = q < nres;
  const int jc = rlist[qv ? q : 0];
  const int kq = l >> 4;
  double bv[21];
#pragma unroll
  for (int s = 0; s < 21; s++) bv[s] = X[(uint32_t)(sgc[4 * s + kq] * B + jc)];
  f64x4 acc[NMT];
#pragma unroll
  for (int i = 0; i < NMT; i++) acc[i] = f64x4{0.0, 0.0, 0.0, 0.0};
  double r80 = 0.0;
#pragma unroll
  for (int s = 0; s < 21; s++) {
    const double* kc = Kf + (4 * s + kq) * B64_LDK + (l & 15);
#pragma unroll
    for (int i = 0; i < NMT; i++)
      acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(kc[i * 16], bv[s], acc[i], 0, 0, 0);
    r80 = r80 + Kf[(4 * s + kq) * B64_LDK + 80] * bv[s];
  }
  // the four k-groups of the column: lanes p, p+16, p+32, p+48
  r80 = r80 + __shfl_xor(r80, 16);
  r80 = r80 + __shfl_xor(r80, 32);
  if (!qv) return;
#pragma unroll
  for (int i = 0; i < NMT; i++)
#pragma unroll
    for (int v = 0; v < 4; v++) Y[(e * 81 + i * 16 + 4 * v + kq) * B + q] = acc[i][v];
  if (kq == 0) Y[(e * 81 + 80) * B + q] = r80;
}

__global__ __launch_bounds__(256) void k_bmv64(int32_t nelmt, int B,
                                               const double* __restrict__ KP,
                                               const int32_t* __restrict__ gdof,
                                               const double* __restrict__ X,
                                               double* __restrict__ Y,
                                               const int32_t* __restrict__ rlist) {
  constexpr int KPT = (3321 + 255) / 256;
  __shared__ double Kf[BMV_KC * B64_LDK];
  __shared__ int32_t sg[2][BMV_KC];
  const int nres = __builtin_amdgcn_readfirstlane(rlist[64]);
  if (nres == 0) return;  // the common case: no column needs a residual
  const int t = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int l = t & 63;
  for (int q = t; q < 3 * 81; q += 256) Kf[(81 + q / 81) * B64_LDK + q % 81] = 0.0;
  if (t >= 81 && t < BMV_KC) { sg[0][t] = 0; sg[1][t] = 0; }
  uint32_t kaddr[KPT];
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const int q = t + 256 * i;
    const int c = q < 3321 ? tri_col(q) : 0;
    const int r = q < 3321 ? q - c * (c + 1) / 2 : 0;
    kaddr[i] = (uint32_t)(c * B64_LDK + r) | ((uint32_t)(r * B64_LDK + c) << 16);
  }
  int64_t e = blockIdx.x;
  double kreg[KPT];
  int32_t greg = 0;
  if (e < nelmt) {
    const double* src = KP + e * SGEO_KPACK;
#pragma unroll
    for (int i = 0; i < KPT; i++) {
      const int q = t + 256 * i;
      kreg[i] = q < 3321 ? src[q] : 0.0;
    }
    if (t < 81) greg = gdof[e * 81 + t];
  }
  if (t < 81) sg[0][t] = greg;
  int cur = 0;
  for (; e < nelmt; e += gridDim.x) {
#pragma unroll
    for (int i = 0; i < KPT; i++) {
      if (t + 256 * i < 3321) {
        Kf[kaddr[i] & 0xffffu] = kreg[i];
        Kf[kaddr[i] >> 16] = kreg[i];
      }
    }
    __syncthreads();  // K_e and its gdof visible
    const int64_t en = e + gridDim.x;
    if (en < nelmt) {  // next element's K and gdof, in flight during the MFMAs
      const double* src = KP + en * SGEO_KPACK;
#pragma unroll
      for (int i = 0; i < KPT; i++) {
        const int q = t + 256 * i;
        kreg[i] = q < 3321 ? src[q] : 0.0;
      }
      if (t < 81) greg = gdof[en * 81 + t];
    }
    bmv64_element(e, nres, B, X, Y, rlist, Kf, sg[cur], w, l);
    if (t < 81) sg[cur ^ 1][t] = greg;
    __syncthreads();  // Kf free for the next element, its gdof visible
    cur ^= 1;
  }
}

// --------------------------------------------------------------------------
// batched vector kernels.  Equations are split into fixed chunks of ck
// equations (ck depends on the equation count only, see batch_chunk);
// thread t owns column j = t % B and chunk blockIdx*(256/B) + t/B, and walks
// its chunk in order, so every per-(chunk, column) partial is a sequential
// sum whose value does not depend on B.  The finalize then combines the
// chunk partials in a fixed order.
// Partials: part[(kind*nch + chunk)*B + j].
// --------------------------------------------------------------------------
struct ChunkIdx {
  int j;
  int64_t chunk, g0, g1;
  bool valid;
};

__device__ __forceinline__ ChunkIdx chunk_of(int64_t n1, int B, int ck) {
  ChunkIdx c;
  const int t = threadIdx.x;
  c.j = t % B;
  c.chunk = (int64_t)blockIdx.x * (blockDim.x / B) + t / B;
  c.g0 = c.chunk * ck;
  c.g1 = c.g0 + ck < n1 ? c.g0 + ck : n1;
  c.valid = c.g0 < n1;
  return c;
}

__device__ __forceinline__ int64_t nchunks(int64_t n1, int ck) {
  return (n1 + ck - 1) / ck;
}

// Small systems (`grouped`: fewer than 16,384 chunks) combine the partials
// of four consecutive chunks inside the block, in chunk order, and write
// one value per group, part[(kind * ng4 + chunk / 4) * B + j] with
// ng4 = ceil(nch / 4), which the final stage reads directly: one launch per
// reduction.  The groups depend on the chunk index only (a block holds
// 256 / B chunks, a multiple of 4), so a column's sums still do not depend
// on B.  Every thread of the block takes part (there are barriers inside);
// `put` says whether the thread's column writes.  Large systems write one
// partial per chunk, part[(kind * nch + chunk) * B + j].
template <bool GROUPED, int N>
__device__ __forceinline__ void put_partials(double* __restrict__ part, double* lds,
                                             int64_t nch, int B, const ChunkIdx& c,
                                             const int (&kind)[N], const double (&v)[N],
                                             const bool (&is_max)[N],
                                             const bool (&put)[N]) {
  if constexpr (!GROUPED) {
#pragma unroll
    for (int n = 0; n < N; n++)
      if (put[n]) part[(kind[n] * nch + c.chunk) * B + c.j] = v[n];
    return;
  }
  // lds holds N x 256 values; one barrier, the kernel's last use of lds
  const int t = threadIdx.x;
#pragma unroll
  for (int n = 0; n < N; n++) lds[n * 256 + t] = v[n];
  __syncthreads();
  if (((t / B) & 3) != 0) return;
  const int64_t ng4 = (nch + 3) / 4;
#pragma unroll
  for (int n = 0; n < N; n++) {
    if (!put[n]) continue;
    double a = v[n];
    for (int k = 1; k < 4 && c.chunk + k < nch; k++) {
      const double w = lds[n * 256 + t + k * B];
      a = is_max[n] ? fmax(a, w) : a + w;
    }
    part[(kind[n] * ng4 + c.chunk / 4) * B + c.j] = a;
  }
}

// p_i = D r + beta p_{i-1} for INNER columns with a pending beta (i the
// iteration about to run; p_i in buffer i & pmask of the n1 x B buffers),
// and the chunk partial of max|p_i| for the iteration's stop test
template <bool GROUPED>
__global__ __launch_bounds__(256) void k_bpupdate(int64_t n1, int B, int ck,
                                                  const float* __restrict__ r,
                                                  const float* __restrict__ dinv,
                                                  float* __restrict__ p,
                                                  double* __restrict__ part,
                                                  const int32_t* __restrict__ cst,
                                                  const BCol* __restrict__ bc) {
  __shared__ double lds[GROUPED ? 256 : 1];
  const ChunkIdx c = chunk_of(n1, B, ck);
  const bool act = c.valid && cst[c.j] == BCOL_INNER && bc[c.j].pending;
  double mp = 0.0;
  if (act) {
    const float beta = (float)bc[c.j].beta;
    const int i = bc[c.j].it + 1;
    const int64_t nv = n1 * B;
    const int pm = bc[c.j].pmask;
    const float* __restrict__ pold = p + ((i - 1) & pm) * nv;
    float* __restrict__ pnew = p + (i & pm) * nv;
    for (int64_t g = c.g0; g < c.g1; g++) {
      const int64_t q = g * B + c.j;
      const float pg = dinv[g] * r[q] + beta * pold[q];
      pnew[q] = pg;
      mp = fmax(mp, fabs((double)pg));
    }
  }
  // max|p_i| (INNER kind 2)
  put_partials<GROUPED, 1>(part, lds, nchunks(n1, ck), B, c, {2}, {mp}, {true}, {act});
}

// The deferred x updates of a column whose solve ended between two x
// writes (DONE or FAIL with npend > 0): x += alpha_k p_k for the npend
// iterations k after the last write, in order; k_bxlag_clear then clears
// npend once every chunk is done.
__global__ __launch_bounds__(256) void k_bmaterialize(int64_t n1, int B, int ck,
                                                      double* __restrict__ x,
                                                      const float* __restrict__ p,
                                                      const int32_t* __restrict__ cst,
                                                      const BCol* __restrict__ bc) {
  const ChunkIdx c = chunk_of(n1, B, ck);
  if (!c.valid) return;
  const int st = cst[c.j];
  const BCol& bj = bc[c.j];
  const int np = bj.npend;  // at most 7
  if ((st != BCOL_DONE && st != BCOL_FAIL) || np == 0) return;
  const int64_t nv = n1 * B;
  double al[8];
  const float* pk[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int it = bj.it - np + 1 + (k < np ? k : 0);
    al[k] = bj.alpha_hist[it & 7];
    pk[k] = p + (it & bj.pmask) * nv;
  }
  for (int64_t g = c.g0; g < c.g1; g++) {
    const int64_t q = g * B + c.j;
    double xg = 0.0;
    if (g != 0) {
      xg = x[q];
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (k < np) xg = xg + al[k] * (double)pk[k][q];
    }
    x[q] = xg;
  }
}

__global__ void k_bxlag_clear(int B, const int32_t* __restrict__ cst,
                              BCol* __restrict__ bc) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= B) return;
  if (cst[j] == BCOL_DONE || cst[j] == BCOL_FAIL) bc[j].npend = 0;
}

// The update of an iteration, with kp assembled on the fly:
//   kp[g][j] = sum over the slots of Y (element order, FP32),
//   r -= alpha kp, z = D r,
// and the r.z partial (FP64) for INNER columns.  x is written by k_bxfold
// on the steps that write it; the max|p| partial of the iteration comes from
// where p was formed (k_bpupdate, or k_bresid for p_1), so the update reads
// neither x nor p.  The (at most 8 on hexahedral meshes) element
// contributions of an equation are loaded together, then summed in slot
// order; UPD_EQ equations per trip keep that many gathers in flight.
// WAVE_CHUNK: B = 64, so a wave's 64 lanes are the 64 columns of one chunk;
// the chunk's row pointers and slots are first staged in the wave's LDS rows
// (stage_indices), so each gather waits on the Y load only, not on a chain
// of row-pointer and slot loads before it.
#ifndef UPD_EQ
#define UPD_EQ 4
#endif
#define UPD_SL 512  // staged slots per wave (64 equations, 8 slots each)
// WAVE_CHUNK: the wave's chunk's row pointers rowptr[ga .. g1] (ga = max(g0,
// 1): equation 0, the sink, has no update) and its slots, one coalesced load
// each, into the wave's LDS rows rp[65], sl[UPD_SL], so the update's gathers
// wait on one memory level instead of three.  All 64 lanes call it; false
// when the slots do not fit (the update then reads the indices from memory).
__device__ __forceinline__ bool stage_indices(const ChunkIdx& c,
                                              const int32_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ slots,
                                              int32_t* rp, int32_t* sl) {
  const int ln = threadIdx.x & 63;
  const int64_t ga = c.g0 > 0 ? c.g0 : 1;
  const int nr = (int)(c.g1 - ga) + 1;  // <= 65
  if (ln < nr) rp[ln] = rowptr[ga + ln];
  if (ln == 0 && nr > 64) rp[64] = rowptr[ga + 64];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int32_t sb = rp[0], nsl = rp[nr - 1] - sb;
  if (nsl > UPD_SL) return false;
  for (int k = ln; k < nsl; k += 64) sl[k] = slots[sb + k];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return true;
}

// the body of k_bupdate for one active (chunk, column); rp / sl: the staged
// indices (null: read them from memory)
template <bool WAVE_CHUNK, bool STAGED>
__device__ __forceinline__ void update_chunk(
    int64_t n1, int B, int ck, const ChunkIdx& c, const int32_t* __restrict__ rowptr,
    const int32_t* __restrict__ slots, const float* __restrict__ Y,
    const float* __restrict__ dinv, float* __restrict__ r, const BCol* __restrict__ bc,
    double& s, const int32_t* rp, const int32_t* sl) {
  const float alpha32 = (float)bc[c.j].alpha;
  const int64_t ga = c.g0 > 0 ? c.g0 : 1;
  const int32_t sb = STAGED ? rp[0] : 0;
  auto gather = [&](int64_t g, float* v, int32_t& q0, int32_t& q1) {
    if constexpr (STAGED) {
      q0 = g > 0 ? rp[g - ga] : 0;
      q1 = g > 0 ? rp[g - ga + 1] : 0;
#pragma unroll
      for (int k = 0; k < 8; k++)
        v[k] = (q0 + k < q1) ? Y[(int64_t)sl[q0 + k - sb] * B + c.j] : 0.0f;
      return;
    }
    q0 = g > 0 ? rowptr[g] : 0;
    q1 = g > 0 ? rowptr[g + 1] : 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
      v[k] = (q0 + k < q1) ? Y[(int64_t)slots[q0 + k] * B + c.j] : 0.0f;
  };
  auto finish = [&](int64_t g, const float* v, int32_t q0, int32_t q1) {
    float kp = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (q0 + k < q1) kp = kp + v[k];
    for (int32_t q = q0 + 8; q < q1; q++)
      kp = kp + Y[(int64_t)(STAGED ? sl[q - sb] : slots[q]) * B + c.j];
    const int64_t q = g * B + c.j;
    const float rg = r[q] - alpha32 * kp;
    r[q] = rg;
    const float zg = dinv[g] * rg;
    s = s + (double)rg * (double)zg;
  };
  int64_t g = c.g0;
  for (; g + UPD_EQ <= c.g1; g += UPD_EQ) {
    float v[UPD_EQ][8];
    int32_t q0[UPD_EQ], q1[UPD_EQ];
#pragma unroll
    for (int k = 0; k < UPD_EQ; k++) gather(g + k, v[k], q0[k], q1[k]);
#pragma unroll
    for (int k = 0; k < UPD_EQ; k++) finish(g + k, v[k], q0[k], q1[k]);
  }
  for (; g < c.g1; g++) {
    float va[8];
    int32_t a0, a1;
    gather(g, va, a0, a1);
    finish(g, va, a0, a1);
  }
}

// one thread per (chunk, column); the r.z partial (kind 1)
template <bool WAVE_CHUNK, bool GROUPED>
__global__ __launch_bounds__(256) void k_bupdate(
    int64_t n1, int B, int ck, const int32_t* __restrict__ rowptr,
    const int32_t* __restrict__ slots, const float* __restrict__ Y,
    const float* __restrict__ dinv, float* __restrict__ r, double* __restrict__ part,
    const int32_t* __restrict__ cst, const BCol* __restrict__ bc) {
  __shared__ double lds[GROUPED ? 256 : 1];
  __shared__ int32_t idx[WAVE_CHUNK ? 4 * (65 + UPD_SL) : 1];
  ChunkIdx c = chunk_of(n1, B, ck);
  if (WAVE_CHUNK) {
    const int ch = __builtin_amdgcn_readfirstlane((int)c.chunk);
    c.chunk = ch;
    c.g0 = (int64_t)ch * ck;
    c.g1 = c.g0 + ck < n1 ? c.g0 + ck : n1;
    c.valid = c.g0 < n1;
  }
  const bool act = c.valid && cst[c.j] == BCOL_INNER;
  const int64_t nch = nchunks(n1, ck);
  const int32_t* rp = nullptr;
  const int32_t* sl = nullptr;
  if constexpr (WAVE_CHUNK) {
    int32_t* wr = idx + (threadIdx.x >> 6) * (65 + UPD_SL);
    // c.valid is uniform over the wave (one chunk per wave); skip when no
    // lane's column iterates
    if (c.valid && __any(act) && stage_indices(c, rowptr, slots, wr, wr + 65)) {
      rp = wr;
      sl = wr + 65;
    }
  }
  double s = 0.0;
  if (act) {
    if (rp)
      update_chunk<WAVE_CHUNK, true>(n1, B, ck, c, rowptr, slots, Y, dinv, r, bc, s, rp, sl);
    else
      update_chunk<WAVE_CHUNK, false>(n1, B, ck, c, rowptr, slots, Y, dinv, r, bc, s, rp, sl);
  }
  put_partials<GROUPED, 1>(part, lds, nch, B, c, {1}, {s}, {false}, {act});
}

// The x write of an iteration i with i % xlen == 0 (INNER columns), after
// the update of i: the npend deferred updates before i and i's own, in
// order, x += alpha_k p_k for k = i - np + 1 .. i (np = npend + 1 <= 8),
// and the max|x| partial (kind 0; x(0) = 0) for i's stop test.
template <bool GROUPED>
__global__ __launch_bounds__(256) void k_bxfold(int64_t n1, int B, int ck,
                                                double* __restrict__ x,
                                                const float* __restrict__ p,
                                                double* __restrict__ part,
                                                const int32_t* __restrict__ cst,
                                                const BCol* __restrict__ bc) {
  __shared__ double lds[GROUPED ? 256 : 1];
  const ChunkIdx c = chunk_of(n1, B, ck);
  bool act = c.valid && cst[c.j] == BCOL_INNER;
  double mx = 0.0;
  if (act) {
    const BCol& bj = bc[c.j];
    const int i = bj.it + 1;
    // a column off the x-writing phase writes nothing here; the INNER
    // finalize reports it
    act = (i & (bj.xlen - 1)) == 0;
    const int np = bj.npend + 1;
    const int64_t nv = n1 * B;
    double al[8];
    uint32_t pk[8];  // buffer offsets, pmask n1 B < 2^32
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int it = i - np + 1 + (k < np ? k : 0);
      al[k] = bj.alpha_hist[it & 7];
      pk[k] = (uint32_t)(it & bj.pmask) * (uint32_t)nv;
    }
    for (int64_t g = act ? c.g0 : c.g1; g < c.g1; g++) {
      const int64_t q = g * B + c.j;
      double xg = 0.0;
      if (g != 0) {
        xg = x[q];
#pragma unroll
        for (int k = 0; k < 8; k++)
          if (k < np) xg = xg + al[k] * (double)p[pk[k] + q];
      }
      x[q] = xg;
      mx = fmax(mx, fabs(xg));
    }
  }
  put_partials<GROUPED, 1>(part, lds, nchunks(n1, ck), B, c, {0}, {mx}, {true}, {act});
}

// At a residual step: INNER columns due for a replacement become REPLACE
// (their last iteration left r_{k+1}, rz_prev = r_k.z_k and a pending beta,
// which the FP64 residual's finalize recomputes), and the columns that need
// an FP64 residual now (START, REPLACE) form the residual list, in column
// order: rlist[0..nres-1], rlist[64] = nres, rslot[j] = column j's place.
// One wave (B <= 64).
__global__ __launch_bounds__(64) void k_bpromote(int B, int32_t* __restrict__ cst,
                                                 BCol* __restrict__ bc,
                                                 int32_t* __restrict__ rlist,
                                                 int32_t* __restrict__ rslot) {
  const int j = threadIdx.x;
  bool want = false;
  if (j < B) {
    int st = cst[j];
    if (st == BCOL_INNER && bc[j].want_rep) {
      st = BCOL_REPLACE;
      cst[j] = st;
      bc[j].want_rep = 0;
    }
    want = st == BCOL_START || st == BCOL_REPLACE;
  }
  const unsigned long long m = __ballot(want);
  const int slot = __popcll(m & ((1ull << j) - 1ull));
  if (want) {
    rlist[slot] = j;
    rslot[j] = slot;
  }
  if (j == 0) rlist[64] = __popcll(m);
}

// FP64 residual for START / REPLACE columns: r64 = load - sum Y64 (element
// order, FP64; Y64 holds the residual list's columns at their list places);
// r = (float) r64; START: p_1 = D r; partials r.(D r) and (START) max|x_0|,
// and for START the first iteration's max|p_1|
template <bool GROUPED>
__global__ __launch_bounds__(256) void k_bresid(
    int64_t n1, int B, int ck, const int32_t* __restrict__ rowptr,
    const int32_t* __restrict__ slots, const double* __restrict__ Y64,
    const double* __restrict__ load, const float* __restrict__ dinv,
    const double* __restrict__ x, float* __restrict__ r, float* __restrict__ p,
    double* __restrict__ part, const int32_t* __restrict__ cst,
    const int32_t* __restrict__ rslot) {
  __shared__ double lds[GROUPED ? 3 * 256 : 1];
  const ChunkIdx c = chunk_of(n1, B, ck);
  const int st = c.valid ? cst[c.j] : BCOL_OFF;
  const bool act = st == BCOL_START || st == BCOL_REPLACE;
  const int qs = act ? rslot[c.j] : 0;
  const int64_t nv = n1 * B;
  double s = 0.0, mx = 0.0, mp = 0.0;
  for (int64_t g = act ? c.g0 : c.g1; g < c.g1; g++) {
    const int64_t q = g * B + c.j;
    double kx = 0.0;
    if (g > 0)
      for (int32_t w = rowptr[g]; w < rowptr[g + 1]; w++)
        kx = kx + Y64[(int64_t)slots[w] * B + qs];
    const double r64 = (g == 0) ? 0.0 : load[q] - kx;
    const float rg = (float)r64;
    const float zg = dinv[g] * rg;
    r[q] = rg;
    if (st == BCOL_START) {  // p_1 = D r (buffer 1); max|x_0|, max|p_1|
      p[nv + q] = zg;
      mx = fmax(mx, fabs(x[q]));
      mp = fmax(mp, fabs((double)zg));
    }
    s = s + (double)rg * (double)zg;
  }
  const int64_t nch = nchunks(n1, ck);
  // kind 2 (max|p_1|, START columns) is the INNER reduction's third kind
  put_partials<GROUPED, 3>(part, lds, nch, B, c, {0, 1, 2}, {s, mx, mp},
                           {false, true, true}, {act, act, act && st == BCOL_START});
}

// Reductions of the chunk partials, in a fixed order independent of B.  On
// large systems stage 1 (k_bgroup, thread per (group, column), column
// fastest: coalesced) combines groups of G consecutive chunks per column
// (batch_group: 64, or 16 for ALPHA below 16,384 elements) and stage 2
// (k_bfinalize, one block per column) combines the groups and advances the
// column's state machine.  Small systems skip
// stage 1: their vector kernels already wrote one partial per four chunks
// (put_partials), and the ALPHA reduction of up to BFIN_DIRECT_ELEMENTS
// elements reads the per-element shares directly.
//   RESID: chunks of equations; r.z of the FP64 residual, max|x| (START).
//   ALPHA: the elements' p.Kp shares from the MFMA product (a "chunk" is an
//          element).
//   INNER: chunks of equations; max|x|, r.z, max|p|.
#define BGROUP_MIN 16
#define BFIN_KINDS 3

__device__ __forceinline__ int fin_states(int what) {
  return what == BFIN_RESID ? (BCOL_BIT(BCOL_START) | BCOL_BIT(BCOL_REPLACE))
                            : BCOL_BIT(BCOL_INNER);
}

__device__ __forceinline__ int fin_kinds(int what) {
  return what == BFIN_INNER ? 3 : what == BFIN_RESID ? 2 : 1;
}

// kind k of `what` combines by max (else by sum)
__device__ __forceinline__ bool fin_max(int what, int k) {
  return (what == BFIN_INNER && k != 1) || (what == BFIN_RESID && k == 1);
}

__global__ __launch_bounds__(256) void k_bgroup(int64_t nch, int B, int G,
                                                const double* __restrict__ part,
                                                double* __restrict__ part2,
                                                const int32_t* __restrict__ cst,
                                                int what) {
  const int64_t ng = (nch + G - 1) / G;
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= ng * B) return;
  const int j = (int)(q % B);
  const int64_t g = q / B;
  if (!wanted(cst, j, fin_states(what))) return;
  const int nk = fin_kinds(what);
  const int64_t b0 = g * G, b1 = b0 + G < nch ? b0 + G : nch;
  for (int k = 0; k < nk; k++) {
    const bool mx = fin_max(what, k);
    const double* src = part + k * nch * B;
    double v = 0.0;
    int64_t b = b0;
    // sixteen partials loaded together, then combined in chunk order
    for (; b + 16 <= b1; b += 16) {
      double a[16];
#pragma unroll
      for (int i = 0; i < 16; i++) a[i] = src[(b + i) * B + j];
#pragma unroll
      for (int i = 0; i < 16; i++) v = mx ? fmax(v, a[i]) : v + a[i];
    }
    for (; b < b1; b++) {
      const double a = src[b * B + j];
      v = mx ? fmax(v, a) : v + a;
    }
    part2[(k * ng + g) * B + j] = v;
  }
}

// `ng` values per column and kind in gp[(kind * ng + b) * B + j]: the
// group partials of the two-stage form, or the partials written directly
// by the kernels (grouped small systems; the element shares of ALPHA on
// small meshes)
__global__ __launch_bounds__(256) void k_bfinalize(int64_t ng, int B,
                                                   const double* __restrict__ gp,
                                                   int32_t* __restrict__ cst,
                                                   BCol* __restrict__ bc,
                                                   int what, int xw) {
  __shared__ double red[BFIN_KINDS * 4];
  const int j = blockIdx.x;
  const int st = cst[j];
  if (!wanted(cst, j, fin_states(what))) return;  // uniform per block
  const int nk = fin_kinds(what);
  // every kind in one pass (their loads overlap); per kind, each thread's
  // values in group order, then a wave tree, then the four waves in order
  double a[BFIN_KINDS] = {0.0, 0.0, 0.0};
  for (int64_t b = threadIdx.x; b < ng; b += 256) {
#pragma unroll
    for (int k = 0; k < BFIN_KINDS; k++) {
      if (k < nk) {
        const double y = gp[(k * ng + b) * B + j];
        a[k] = fin_max(what, k) ? fmax(a[k], y) : a[k] + y;
      }
    }
  }
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < BFIN_KINDS; k++) {
    if (k < nk) {
      const double t = fin_max(what, k) ? wave_max(a[k]) : wave_sum(a[k]);
      if ((threadIdx.x & 63) == 0) red[k * 4 + w] = t;
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double v[BFIN_KINDS] = {0.0, 0.0, 0.0};
  for (int k = 0; k < nk; k++) {
    const bool mx = fin_max(what, k);
#pragma unroll
    for (int i = 0; i < 4; i++) v[k] = mx ? fmax(v[k], red[k * 4 + i]) : v[k] + red[k * 4 + i];
  }
  BCol& c = bc[j];
  switch (what) {
    case BFIN_RESID:  // v0 = r.z of the FP64 residual, v1 = max|x| (START)
      c.nrep += 1;
      if (st == BCOL_START) {
        c.rz = v[0];
        c.rz0 = v[0];
        c.it = 0;
        c.pending = 0;
        c.maxx = v[1];
        c.xsum = 0.0;
        c.npend = 0;
        cst[j] = (v[0] == 0.0) ? BCOL_DONE : BCOL_INNER;  // r = 0: x is exact
      } else {  // replaces r_{k+1}: beta = r.z (true) / r_k.z_k
        c.beta = v[0] / c.rz_prev;
        c.rz = v[0];
        c.rz0 = v[0];
        c.pending = 1;
        cst[j] = BCOL_INNER;
      }
      break;
    case BFIN_ALPHA:  // v0 = p.Kp
      c.alpha = c.rz / v[0];
      c.alpha_hist[(c.it + 1) & 7] = c.alpha;
      c.pending = 0;
      break;
    case BFIN_INNER: {  // v0 = max|x| (x written), v1 = r.z, v2 = max|p|
      c.it += 1;
      c.maxp = v[2];
      const double a = fabs(c.alpha) * v[2];
      const bool xit = (c.it & (c.xlen - 1)) == 0;
      if (xit != (xw != 0)) {
        // the step's update and the column disagree on whether x was
        // written (the engine's step phase and the column's iteration
        // count went apart): the column ends, and the host reports it
        c.xerr = 1;
        cst[j] = BCOL_FAIL;
        break;
      }
      bool stop;
      if (xit) {
        // relative update test of src/solver.f90:65 (after the x update)
        stop = a / v[0] <= c.tol;
        c.maxx = v[0];
        c.xsum = 0.0;
        c.npend = 0;
      } else {
        // x_it = x_w + sum of alpha_k p_k over the iterations after the last
        // write w is not written yet, but max|x_it| >= max|x_w| - xsum with
        // xsum the sum of |alpha_k| max|p_k|: a <= tol (max|x_w| - xsum)
        // passes the same test for certain.  Only inside the window of
        // relative width ~tol xsum / max|x| (~xlen tol^2 near convergence)
        // where the bound cannot decide does the solve iterate once more
        // than the test would.
        c.xsum += a;
        stop = a <= c.tol * (c.maxx - c.xsum);
        c.npend += 1;
      }
      if (stop) {
        cst[j] = BCOL_DONE;
      } else if (c.it >= c.maxiter) {
        cst[j] = BCOL_FAIL;
      } else {
        c.rz_prev = c.rz;
        // due for an FP64 residual: replaced at the next residual step, by
        // when the column may have iterated on
        if (v[1] <= c.inner_tol * c.rz0) c.want_rep = 1;
        c.beta = v[1] / c.rz;
        c.rz = v[1];
        c.pending = 1;
      }
      break;
    }
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
__global__ __launch_bounds__(256) void k_buerr(int64_t n1, int B, int ck,
                                               double* __restrict__ x,
                                               double* __restrict__ oldx,
                                               double* __restrict__ part,
                                               const int32_t* __restrict__ active) {
  const ChunkIdx c = chunk_of(n1, B, ck);
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
  const int64_t nch = nchunks(n1, ck);
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
#pragma unroll 4
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

// Chunk length of the vector kernels: a function of the equation count
// only (never of B), so a column's partial sums stay independent of the
// batch.  Large systems use 64 equations per thread; small ones shorter
// chunks, so that the launch still has enough waves to hide the latency of
// each thread's walk.
static int batch_chunk(int64_t n1) {
  int ck = 64;
  while (ck > 4 && n1 / ck < 16384) ck /= 2;
  return ck;
}

static inline int64_t nchunks_host(int64_t n1) {
  const int ck = batch_chunk(n1);
  return (n1 + ck - 1) / ck;
}

// stage-1 group size of the reductions (a function of the chunk count only)
static inline int batch_group(int64_t nch) { return nch < 16384 ? 16 : 64; }
// small systems: the vector kernels combine groups of four chunks themselves
static inline bool batch_grouped(int64_t nch) { return batch_group(nch) == BGROUP_MIN; }

int64_t batch_partials(int64_t n1) {
  const int64_t nch = nchunks_host(n1);
  // chunk + group partials of every kind
  return BFIN_KINDS * (nch + (nch + BGROUP_MIN - 1) / BGROUP_MIN);
}

int64_t batch_element_partials(int32_t nelmt) {
  return (int64_t)nelmt + ((int64_t)nelmt + BGROUP_MIN - 1) / BGROUP_MIN;
}

int64_t batch_equation_chunks(int64_t n1) { return nchunks_host(n1); }

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

// Launch geometry of the persistent MFMA product: blocks per CU (at most 3:
// the kernel is built for 3 waves per SIMD, at most 168 VGPRs) and the
// XCD-aware element split (nxcd 8) or a plain grid-stride walk (nxcd 1).
void launch_bmv_mfma_f32(hipStream_t s, int32_t nelmt, int B, int n_cu, BmvGeom geom,
                         const float* KP, const int32_t* gdof, const int32_t* eorder,
                         const float* P, int64_t nv, float* Y, double* edot,
                         const int32_t* cst, const BCol* bc, int want_mask) {
  const int nx = geom.xcd_split ? BMV_XCD : 1;
  const int bpc = geom.blocks_per_cu < 1 ? 1 : geom.blocks_per_cu > 4 ? 4 : geom.blocks_per_cu;
  int64_t g = (int64_t)n_cu * bpc;
  const int64_t need = ((int64_t)nelmt + nx - 1) / nx * nx;
  if (g > need) g = need;
  g = (g + nx - 1) / nx * nx;  // a multiple of the XCD count
  const dim3 grid((unsigned)g);
  if (B >= 64)
    hipLaunchKernelGGL(k_bmv_mfma<4>, grid, dim3(256), 0, s, nelmt, B, nx, KP, gdof,
                       eorder, P, nv, Y, edot, cst, bc, want_mask);
  else if (B == 32)
    hipLaunchKernelGGL(k_bmv_mfma<2>, grid, dim3(256), 0, s, nelmt, B, nx, KP, gdof,
                       eorder, P, nv, Y, edot, cst, bc, want_mask);
  else
    hipLaunchKernelGGL(k_bmv_mfma<1>, grid, dim3(256), 0, s, nelmt, B, nx, KP, gdof,
                       eorder, P, nv, Y, edot, cst, bc, want_mask);
}

// FP64 residual product: persistent, 2 blocks per CU (60 KB of LDS each);
// every block returns at once when the residual list is empty
void launch_bmv_f64(hipStream_t s, int32_t nelmt, int B, int n_cu, const double* KP,
                    const int32_t* gdof, const double* X, double* Y,
                    const int32_t* rlist) {
  int64_t g = (int64_t)n_cu * 2;
  if (g > nelmt) g = nelmt;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_bmv64, dim3((unsigned)g), dim3(256), 0, s, nelmt, B, KP, gdof, X,
                     Y, rlist);
}

void launch_bupdate(hipStream_t s, int64_t n1, int B, const int32_t* rowptr,
                    const int32_t* slots, const float* Y, const float* dinv, float* r,
                    double* part, const int32_t* cst, const BCol* bc) {
  const int ck = batch_chunk(n1);
  const bool gr = batch_grouped(nchunks_host(n1));
#define SGEO_UPD(W, G)                                                                  \
  hipLaunchKernelGGL((k_bupdate<W, G>), dim3(chunk_blocks(n1, B)), dim3(256), 0, s, n1, \
                     B, ck, rowptr, slots, Y, dinv, r, part, cst, bc)
  if (B == 64) {
    if (gr) SGEO_UPD(true, true); else SGEO_UPD(true, false);
  } else {
    if (gr) SGEO_UPD(false, true); else SGEO_UPD(false, false);
  }
#undef SGEO_UPD
}

void launch_bxfold(hipStream_t s, int64_t n1, int B, double* x, const float* p,
                   double* part, const int32_t* cst, const BCol* bc) {
  if (batch_grouped(nchunks_host(n1)))
    hipLaunchKernelGGL(k_bxfold<true>, dim3(chunk_blocks(n1, B)), dim3(256), 0, s, n1, B,
                       batch_chunk(n1), x, p, part, cst, bc);
  else
    hipLaunchKernelGGL(k_bxfold<false>, dim3(chunk_blocks(n1, B)), dim3(256), 0, s, n1, B,
                       batch_chunk(n1), x, p, part, cst, bc);
}

void launch_bpupdate(hipStream_t s, int64_t n1, int B, const float* r,
                     const float* dinv, float* p, double* part, const int32_t* cst,
                     const BCol* bc) {
  if (batch_grouped(nchunks_host(n1)))
    hipLaunchKernelGGL(k_bpupdate<true>, dim3(chunk_blocks(n1, B)), dim3(256), 0, s, n1, B,
                       batch_chunk(n1), r, dinv, p, part, cst, bc);
  else
    hipLaunchKernelGGL(k_bpupdate<false>, dim3(chunk_blocks(n1, B)), dim3(256), 0, s, n1,
                       B, batch_chunk(n1), r, dinv, p, part, cst, bc);
}

void launch_bpromote(hipStream_t s, int B, int32_t* cst, BCol* bc, int32_t* rlist,
                     int32_t* rslot) {
  hipLaunchKernelGGL(k_bpromote, dim3(1), dim3(64), 0, s, B, cst, bc, rlist, rslot);
}

void launch_bresid(hipStream_t s, int64_t n1, int B, const int32_t* rowptr,
                   const int32_t* slots, const double* Y64, const double* load,
                   const float* dinv, const double* x, float* r, float* p,
                   double* part, const int32_t* cst, const int32_t* rslot) {
  if (batch_grouped(nchunks_host(n1)))
    hipLaunchKernelGGL(k_bresid<true>, dim3(chunk_blocks(n1, B)), dim3(256), 0, s, n1, B,
                       batch_chunk(n1), rowptr, slots, Y64, load, dinv, x, r, p, part, cst,
                       rslot);
  else
    hipLaunchKernelGGL(k_bresid<false>, dim3(chunk_blocks(n1, B)), dim3(256), 0, s, n1, B,
                       batch_chunk(n1), rowptr, slots, Y64, load, dinv, x, r, p, part, cst,
                       rslot);
}

void launch_bmaterialize(hipStream_t s, int64_t n1, int B, double* x, const float* p,
                         const int32_t* cst, BCol* bc) {
  hipLaunchKernelGGL(k_bmaterialize, dim3(chunk_blocks(n1, B)), dim3(256), 0, s, n1, B,
                     batch_chunk(n1), x, p, cst, bc);
  hipLaunchKernelGGL(k_bxlag_clear, dim3((B + 63) / 64), dim3(64), 0, s, B, cst, bc);
}

#define BFIN_DIRECT_ELEMENTS 4096

// `nch` partials per column and kind in `part` (equation chunks, or elements
// for ALPHA); the group partials follow them.  Small systems: one launch,
// the final stage reading what the kernels wrote (groups of four chunks, or
// for ALPHA the element shares of up to BFIN_DIRECT_ELEMENTS elements).
// Large ones: the coalesced group stage, then the final stage.
void launch_bfinalize(hipStream_t s, int64_t nch, int B, double* part,
                      int32_t* cst, BCol* bc, int what, int xw) {
  const bool direct = what == BFIN_ALPHA ? nch <= BFIN_DIRECT_ELEMENTS : batch_grouped(nch);
  if (direct) {
    const int64_t ng = what == BFIN_ALPHA ? nch : (nch + 3) / 4;
    hipLaunchKernelGGL(k_bfinalize, dim3(B), dim3(256), 0, s, ng, B, part, cst, bc, what,
                       xw);
    return;
  }
  const int G = batch_group(nch);
  const int64_t ng = (nch + G - 1) / G;
  // group partials after the chunk partials: for the equation-chunk
  // reductions after all three kinds (RESID runs before k_bpupdate writes
  // kind 2), for ALPHA after the element shares
  double* part2 = part + (what == BFIN_ALPHA ? 1 : BFIN_KINDS) * nch * B;
  hipLaunchKernelGGL(k_bgroup, dim3((unsigned)((ng * B + 255) / 256)), dim3(256), 0,
                     s, nch, B, G, part, part2, cst, what);
  hipLaunchKernelGGL(k_bfinalize, dim3(B), dim3(256), 0, s, ng, B, part2, cst, bc, what,
                     xw);
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
                     batch_chunk(n1), x, oldx, part, active);
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
