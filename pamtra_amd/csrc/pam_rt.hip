// pam_rt.hip -- the RT4 solver on the device.
//
//  k_doubling: one workgroup per (scattering layer, channel).  Takes the
//    m=0 phase matrix that k_phase (pam_phase.hip) formed in the unit's
//    R/T slots (scatcnv.f90:103-192), checks its normalisation
//    (radscat4.f90:87-136), initialises the thin layer (radintg4.f90:1-73)
//    and doubles it up (radintg4.f90:174-309) on three LDS matrices (see
//    DblLds3); the Gauss-Jordan queue pass keeps the four-matrix step
//    (doubling_unit).  Only the + quadrants are
//    formed: for randomly oriented particles with NSTOKES=2 the (-,-) and
//    (+,-) quadrants equal (+,+) and (-,+) exactly, so R- = R+ and T- = T+
//    throughout the symmetric doubling.
//  k_adding: one workgroup per (column, channel).  Streams the layers top
//    down, combining each into the stack (radintg4.f90:447-513).  Clear
//    layers (R = 0, T diagonal) take a closed form that performs the same
//    floating-point operations as the full combine restricted to the
//    non-zero terms.  Then the surface, the internal radiances at the
//    output levels (radintg4.f90:119-172) and the Planck inversion to TB
//    (collect_output.f90:31-88).
//
// Matrices live in LDS with a padded leading dimension (at NUMMU = 32 the
// adding kernels keep their working matrices in a global scratch per
// workgroup, see add_mat).  Products run on the FP64 matrix cores (16x16
// output tiles, one per wave at N = 32, two at N = 64); the inverse is a
// register-resident Gauss-Jordan with partial pivoting (the reference calls
// LAPACK DGETRF/DGETRI).  The same source serves NUMMU = 16 (N = 32) and
// NUMMU = 32 (N = 64).
#include "pam_internal.h"

namespace pam {

// 256 threads at N = 32; at N = 64, where the LDS admits one workgroup per
// CU, 512 threads, so that every SIMD holds two waves
constexpr int RT_WG = (NN == 32) ? 256 : 512;
constexpr int NWAVE = RT_WG / 64;
constexpr int LD = NN + 1;
constexpr int MSZ = NN * LD;
static_assert(NN == 32 || NN == 64, "the tilings below cover N = 32 and N = 64");
constexpr int NT = NN / 16;            // 16x16 tiles per matrix dimension
constexpr int TPW = NT * NT / NWAVE;   // output tiles per wave (1 at N = 32, 2 at N = 64)
// wave wv's tile q: k = wv + NWAVE q, rows 16 (k / NT), columns 16 (k % NT)
__device__ __forceinline__ int tile_i(int wv, int q) { return ((wv + NWAVE * q) / NT) * 16; }
__device__ __forceinline__ int tile_j(int wv, int q) { return ((wv + NWAVE * q) % NT) * 16; }
// matrix-vector mappings: VL lanes per row when the workgroup runs four
// vectors at once (4 x NN rows), RPL lanes per row for one vector (NN rows)
constexpr int VL = RT_WG / (4 * NN);
constexpr int RPL = RT_WG / NN;
// unrolling of the k loops of the products: complete at N = 32.  At N = 64,
// where TPW = 2 tiles per wave give independent chains, by 8 in the
// doubling kernels (one workgroup per CU: the loads of the next steps must
// be in flight) and by 2 in the adding kernels (their other registers leave
// no room for more)
constexpr int KUNR = (NN == 32) ? NN / 4 : 2;
constexpr int KUNR_DBL = (NN == 32) ? NN / 4 : 8;
// k-loop unrolling of the hot doubling step: by 4 at N = 32, which keeps it at
// 88 VGPRs, within the five waves per SIMD its LDS allows (by 8 it needs 116)
constexpr int DBL_STEP_UNR = (NN == 32) ? 4 : 8;

// ------------------------------------------------------------------
// LDS dense helpers; every thread of the workgroup calls them
// ------------------------------------------------------------------

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// NaN-propagating maximum (a NaN anywhere wins)
__device__ __forceinline__ double max_nan(double a, double b) { return (b > a || b != b) ? b : a; }

__device__ __forceinline__ float max_nan(float a, float b) { return (b > a || b != b) ? b : a; }

// maximum of the per-wave maxima (NaN propagates)
__device__ __forceinline__ float wmax_all(const float *w) {
  float m = max_nan(max_nan(w[0], w[1]), max_nan(w[2], w[3]));
#pragma unroll
  for (int q = 4; q < NWAVE; q++) m = max_nan(m, w[q]);
  return m;
}

// Wavefront maximum in single precision (the inverse series only compares it
// against thresholds far from the float limits); NaN propagates.
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max_nan(v, __shfl_xor(v, o, 64));
  return v;
}

// C = D + A * B (D may be NULL) on FP64 matrix cores.  Each wave of the
// workgroup owns TPW 16x16 tiles of C and runs NN/4 v_mfma_f64_16x16x4 steps
// over k for each.  Operand layout of the 16x16x4 f64 MFMA: lane l supplies
// A[l%16][l/16] and B[l/16][l%16] of the 16x4 / 4x16 blocks, and holds
// D[4*r+l/16][l%16], r = 0..3, of the result (rows interleaved over the
// four 16-lane groups).  Every lane has read all its operands before the
// first write, so C may alias A, B or D.  With wmax (4 LDS floats) the
// return value is max |C_ij| (NaN if any entry is NaN), else 0.
template <int U = KUNR>
__device__ __forceinline__ double mm_core(const double *A, const double *B, const double *D,
                                          double *C, float *wmax) {
  const int t = threadIdx.x, wv = t >> 6, l = t & 63;
  const int li = l & 15, lk = l >> 4;
  d4_t acc[TPW];
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) acc[q][r] = D ? D[(I + 4 * r + lk) * LD + J + li] : 0.0;
  }
#pragma unroll U
  for (int s = 0; s < NN / 4; s++) {
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int I = tile_i(wv, q), J = tile_j(wv, q);
      const double a = A[(I + li) * LD + 4 * s + lk];
      const double b = B[(4 * s + lk) * LD + J + li];
      acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
    }
  }
  __syncthreads();
  double m = 0.0;
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) C[(I + 4 * r + lk) * LD + J + li] = acc[q][r];
    m = max_nan(m, max_nan(max_nan(fabs(acc[q][0]), fabs(acc[q][1])),
                           max_nan(fabs(acc[q][2]), fabs(acc[q][3]))));
  }
  if (wmax) {
    const float mf = wave_max((float)m);
    if (l == 0) wmax[wv] = mf;
  }
  __syncthreads();
  return wmax ? (double)wmax_all(wmax) : 0.0;
}

// Two independent products in one pass: C1 = D1 + A1 B1, C2 = D2 + A2 B2
// (D1, D2 may be NULL).  Each wave interleaves the two MFMA chains of its
// tiles; all operands are read before the first write, so any output may
// alias any input.  With wmax the return value is max |C2_ij|.
template <int U = KUNR>
__device__ __forceinline__ double mm2(const double *A1, const double *B1, const double *D1,
                                      double *C1, const double *A2, const double *B2,
                                      const double *D2, double *C2, float *wmax) {
  const int t = threadIdx.x, wv = t >> 6, l = t & 63;
  const int li = l & 15, lk = l >> 4;
  d4_t acc1[TPW], acc2[TPW];
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      acc1[q][r] = D1 ? D1[(I + 4 * r + lk) * LD + J + li] : 0.0;
      acc2[q][r] = D2 ? D2[(I + 4 * r + lk) * LD + J + li] : 0.0;
    }
  }
#pragma unroll U
  for (int s = 0; s < NN / 4; s++) {
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int I = tile_i(wv, q), J = tile_j(wv, q);
      const double a1 = A1[(I + li) * LD + 4 * s + lk];
      const double b1 = B1[(4 * s + lk) * LD + J + li];
      const double a2 = A2[(I + li) * LD + 4 * s + lk];
      const double b2 = B2[(4 * s + lk) * LD + J + li];
      acc1[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc1[q], 0, 0, 0);
      acc2[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, b2, acc2[q], 0, 0, 0);
    }
  }
  __syncthreads();
  double m = 0.0;
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      C1[(I + 4 * r + lk) * LD + J + li] = acc1[q][r];
      C2[(I + 4 * r + lk) * LD + J + li] = acc2[q][r];
    }
    m = max_nan(m, max_nan(max_nan(fabs(acc2[q][0]), fabs(acc2[q][1])),
                           max_nan(fabs(acc2[q][2]), fabs(acc2[q][3]))));
  }
  if (wmax) {
    const float mf = wave_max((float)m);
    if (l == 0) wmax[wv] = mf;
  }
  __syncthreads();
  return wmax ? (double)wmax_all(wmax) : 0.0;
}

// C = A * B
template <int U = KUNR>
__device__ __forceinline__ void mm(const double *A, const double *B, double *C) {
  mm_core<U>(A, B, nullptr, C, nullptr);
}

// C = D + A * B
template <int U = KUNR>
__device__ __forceinline__ void mm_add(const double *D, const double *A, const double *B,
                                       double *C) {
  mm_core<U>(A, B, D, C, nullptr);
}

// row i of A x for a 32-vector (no barrier)
__device__ __forceinline__ double mv_row(const double *A, const double *x, int i) {
  double s = 0.0;
#pragma unroll 8
  for (int k = 0; k < NN; k++) s += A[i * LD + k] * x[k];
  return s;
}

// The same helpers in FP32 (the mixed-precision adding), on the FP32 matrix
// cores.  Layout of the 16x16x4 f32 MFMA: lane l supplies A[l%16][l/16] and
// B[l/16][l%16] and holds C[4*(l/16)+r][l%16], r = 0..3.
template <int U = KUNR>
__device__ __forceinline__ float mm_core(const float *A, const float *B, const float *D, float *C,
                                         float *wmax) {
  const int t = threadIdx.x, wv = t >> 6, l = t & 63;
  const int li = l & 15, lk = l >> 4;
  f4_t acc[TPW];
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) acc[q][r] = D ? D[(I + 4 * lk + r) * LD + J + li] : 0.f;
  }
#pragma unroll U
  for (int s = 0; s < NN / 4; s++) {
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int I = tile_i(wv, q), J = tile_j(wv, q);
      const float a = A[(I + li) * LD + 4 * s + lk];
      const float b = B[(4 * s + lk) * LD + J + li];
      acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[q], 0, 0, 0);
    }
  }
  __syncthreads();
  float m = 0.f;
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) C[(I + 4 * lk + r) * LD + J + li] = acc[q][r];
    m = max_nan(m, max_nan(max_nan(fabsf(acc[q][0]), fabsf(acc[q][1])),
                           max_nan(fabsf(acc[q][2]), fabsf(acc[q][3]))));
  }
  if (wmax) {
    const float mf = wave_max(m);
    if (l == 0) wmax[wv] = mf;
  }
  __syncthreads();
  return wmax ? wmax_all(wmax) : 0.f;
}

template <int U = KUNR>
__device__ __forceinline__ float mm2(const float *A1, const float *B1, const float *D1, float *C1,
                                     const float *A2, const float *B2, const float *D2, float *C2,
                                     float *wmax) {
  const int t = threadIdx.x, wv = t >> 6, l = t & 63;
  const int li = l & 15, lk = l >> 4;
  f4_t acc1[TPW], acc2[TPW];
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      acc1[q][r] = D1 ? D1[(I + 4 * lk + r) * LD + J + li] : 0.f;
      acc2[q][r] = D2 ? D2[(I + 4 * lk + r) * LD + J + li] : 0.f;
    }
  }
#pragma unroll U
  for (int s = 0; s < NN / 4; s++) {
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int I = tile_i(wv, q), J = tile_j(wv, q);
      const float a1 = A1[(I + li) * LD + 4 * s + lk], b1 = B1[(4 * s + lk) * LD + J + li];
      const float a2 = A2[(I + li) * LD + 4 * s + lk], b2 = B2[(4 * s + lk) * LD + J + li];
      acc1[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc1[q], 0, 0, 0);
      acc2[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, b2, acc2[q], 0, 0, 0);
    }
  }
  __syncthreads();
  float m = 0.f;
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      C1[(I + 4 * lk + r) * LD + J + li] = acc1[q][r];
      C2[(I + 4 * lk + r) * LD + J + li] = acc2[q][r];
    }
    m = max_nan(m, max_nan(max_nan(fabsf(acc2[q][0]), fabsf(acc2[q][1])),
                           max_nan(fabsf(acc2[q][2]), fabsf(acc2[q][3]))));
  }
  if (wmax) {
    const float mf = wave_max(m);
    if (l == 0) wmax[wv] = mf;
  }
  __syncthreads();
  return wmax ? wmax_all(wmax) : 0.f;
}

template <int U = KUNR>
__device__ __forceinline__ void mm(const float *A, const float *B, float *C) {
  mm_core<U>(A, B, nullptr, C, nullptr);
}

template <int U = KUNR>
__device__ __forceinline__ void mm_add(const float *D, const float *A, const float *B, float *C) {
  mm_core<U>(A, B, D, C, nullptr);
}

__device__ __forceinline__ float mv_row(const float *A, const float *x, int i) {
  float s = 0.f;
#pragma unroll 8
  for (int k = 0; k < NN; k++) s += A[i * LD + k] * x[k];
  return s;
}

// Row i of A x for a 32-vector, split over the two lanes t and t^1 (lane h
// sums the columns 8h..8h+7 and 16+8h..16+8h+7, so the two lanes of a pair
// read LDS banks 16 apart); both lanes return the full sum.  No barrier.
__device__ __forceinline__ double mv_pair(const double *A, const double *x, int i) {
  const int h = (threadIdx.x & 1) * 8;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < NN / 2; k++) {
    const int c = (k & 7) + h + 16 * (k >> 3);
    s += A[i * LD + c] * x[c];
  }
  return s + __shfl_xor(s, 1, 64);
}

// Row i of A x in the four-vector mapping: a lane pair per row at N = 32,
// one lane per row at N = 64 (i = (t % 64) / VL, lead lane t % VL == 0)
__device__ __forceinline__ double mv_vl(const double *A, const double *x, int i) {
  if constexpr (VL == 2)
    return mv_pair(A, x, i);
  else
    return mv_row(A, x, i);
}
// the four vectors are the groups of RT_WG / 4 threads
__device__ __forceinline__ int vl_group() { return threadIdx.x / (RT_WG / 4); }
__device__ __forceinline__ int vl_row() { return (threadIdx.x % (RT_WG / 4)) / VL; }
__device__ __forceinline__ bool vl_lead() { return (threadIdx.x % VL) == 0; }

// LDS scratch of the inverses; gj_only: Gauss-Jordan without trying the
// series first (set by the queue pass in the Gauss-Jordan test mode)
struct InvLds {
  int gj_only;
  float wmax[NWAVE];
  double red[NWAVE];
  int redi[NWAVE];
  double prow[2][2 * NN];
};

// G = inv(I - X) by Gauss-Jordan elimination with partial pivoting on the
// augmented [I - X | I].  Rows stay in registers: thread t owns row t/RPL
// and the GJC augmented columns GJC*(t%RPL)..+GJC-1 (8 of 64 at N = 32, 32
// of 128 at N = 64).  Pivoting is implicit (no row swaps): the pivot row of
// step k becomes row k of the inverse.  Two barriers per pivot.  Returns
// false if a pivot is zero (singular matrix).  Only the Gauss-Jordan kernels
// use it, so the series kernels keep a low register count.
constexpr int GJC = 2 * NN / RPL;
__device__ bool inv_i_minus(const double *X, double *G, InvLds &sh) {
  const int t = threadIdx.x, r = t / RPL, cg = t % RPL, l = t & 63, wv = t >> 6;
  double a[GJC];
#pragma unroll
  for (int c = 0; c < GJC; c++) {
    const int j = cg * GJC + c;
    a[c] = (j < NN) ? ((r == j ? 1.0 : 0.0) - X[r * LD + j]) : ((r == j - NN) ? 1.0 : 0.0);
  }
  bool used = false, ok = true;
  int my_step = -1;
  for (int k = 0; k < NN; k++) {
    const int kg = k / GJC, kc = k % GJC;
    double ck = 0.0;
#pragma unroll
    for (int c = 0; c < GJC; c++)
      if (c == kc) ck = a[c];
    // entry of my row in column k (held by the lane of column group kg)
    const double f = __shfl(ck, (l & ~(RPL - 1)) | kg, 64);
    // pivot: largest |entry| among rows not yet used, lowest row on ties
    double v = (!used && cg == 0) ? fabs(f) : -1.0;
    int idx = r;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      const double v2 = __shfl_xor(v, o, 64);
      const int i2 = __shfl_xor(idx, o, 64);
      if (v2 > v || (v2 == v && i2 < idx)) {
        v = v2;
        idx = i2;
      }
    }
    if (l == 0) {
      sh.red[wv] = v;
      sh.redi[wv] = idx;
    }
    __syncthreads();
    double bv = sh.red[0];
    int p = sh.redi[0];
#pragma unroll
    for (int q = 1; q < NWAVE; q++) {
      const double v2 = sh.red[q];
      const int i2 = sh.redi[q];
      if (v2 > bv || (v2 == bv && i2 < p)) {
        bv = v2;
        p = i2;
      }
    }
    if (!(bv > 0.0)) ok = false;
    double *prow = sh.prow[k & 1];
    if (r == p) {
#pragma unroll
      for (int c = 0; c < GJC; c++) {
        a[c] = a[c] / f;
        prow[cg * GJC + c] = a[c];
      }
      used = true;
      my_step = k;
    }
    __syncthreads();
    if (r != p) {
#pragma unroll
      for (int c = 0; c < GJC; c++) a[c] = a[c] - f * prow[cg * GJC + c];
    }
  }
  if (cg >= RPL / 2 && my_step >= 0) {
#pragma unroll
    for (int c = 0; c < GJC; c++) G[my_step * LD + (cg * GJC + c - NN)] = a[c];
  }
  __syncthreads();
  return ok;
}

// Truncation threshold of the product-form series below: the omitted tail
// X^(2^k) (I - X)^-1 has entries below N * 2^-56 ||G||, under the rounding
// of the factors themselves.
constexpr double NEU_TOL = 0x1p-56;
constexpr int NEU_MAXIT = 20;

// G = inv(I - A B) by the product-form series, X = A B formed in X
// (overwritten): when the spectral radius of X is below one,
//   inv(I - X) = (I + X)(I + X^2)(I + X^4)...,
// each factor costing two matrix-core products; the series stops as soon as
// max |X^(2^k)| < NEU_TOL (a single product for the thin layers at the start
// of the doubling, where X is ~1e-12).  Returns false if it has not
// converged after NEU_MAXIT factors (spectral radius within ~1e-6 of one) or
// X holds a NaN.  np counts the 32x32 products executed.
// In FP32 (T = float, the mixed-precision adding) the series stops at
// 2^-28, below the single-precision resolution of G.
template <typename T> struct SeriesTol;
template <> struct SeriesTol<double> { static constexpr double v = NEU_TOL; };
template <> struct SeriesTol<float> { static constexpr float v = 0x1p-28f; };

template <int U = KUNR, typename T>
__device__ __forceinline__ bool series_inv(const T *A, const T *B, T *X, T *G, InvLds &sh,
                                           int &np) {
  constexpr T tol = SeriesTol<T>::v;
  const int t = threadIdx.x;
  np++;
  T m = mm_core<U>(A, B, nullptr, X, sh.wmax);
  for (int e = t; e < NN2; e += RT_WG) {
    const int i = e / NN, j = e % NN;
    G[i * LD + j] = (i == j ? (T)1.0 : (T)0.0) + X[i * LD + j];
  }
  __syncthreads();
  // G = I + P with P = X; then per factor: G <- G (I + P), P <- P^2, the
  // two products fused while the next factor still matters
  if (!(m < (T)0x1p64)) return false;          // NaN or overflow
  if ((T)NN * m * m < tol) return true;  // bound on max |P^2|
  m = mm_core<U>(X, X, nullptr, X, sh.wmax);
  np++;
  for (int it = 0; it < NEU_MAXIT; it++) {
    if (!(m < (T)0x1p64)) return false;
    if (m < tol) return true;
    if ((T)NN * m * m < tol) {  // last factor
      mm_add<U>(G, G, X, G);
      np++;
      return true;
    }
    m = mm2<U>(G, X, G, G, X, X, nullptr, X, sh.wmax);
    np += 2;
  }
  return false;
}

// G = inv(I - A B).  GJ = false (the hot kernels): the series only; false
// means it did not converge and the caller queues its unit.  GJ = true (the
// queue pass): the series, and Gauss-Jordan with partial pivoting where the
// series does not converge (Gauss-Jordan only when sh.gj_only); false means
// a singular matrix.
template <bool GJ, int U = KUNR, typename T>
__device__ __forceinline__ bool inv_i_minus_ab(const T *A, const T *B, T *X, T *G, InvLds &sh,
                                               int &np) {
  static_assert(!GJ || sizeof(T) == 8, "the Gauss-Jordan pass runs in FP64");
  if constexpr (GJ) {
    if (!sh.gj_only && series_inv<U>(A, B, X, G, sh, np)) return true;
    np++;
    mm<U>(A, B, X);
    return inv_i_minus(X, G, sh);
  } else {
    return series_inv<U>(A, B, X, G, sh, np);
  }
}

// ------------------------------------------------------------------
// phase matrix + doubling
// ------------------------------------------------------------------

struct DblLds {
  double R[MSZ], T[MSZ], G[MSZ], X[MSZ];
  double v[12][NN];  // S+, S-, C+, C-, then source-chain scratch
  double norm[NMU];
  InvLds inv;
  double mu[NMU], qw[NMU];
  int slot;  // unit index of a persistent workgroup
};

// One (scattering layer, channel) unit.  Returns false only in the series
// variant (GJ = false) when an inverse did not converge; nothing has been
// written for the unit then.
template <bool GJ>
__device__ bool doubling_unit(const Dims &d, const Settings &s, const double *freqs,
                              const Work &w, int itf, DblLds &L) {
  double *sR = L.R, *sT = L.T, *sG = L.G, *sX = L.X;
  double(*sv)[NN] = L.v;
  double *snorm = L.norm, *smu = L.mu, *sqw = L.qw;
  InvLds &sinv = L.inv;
  // until INITIALIZE4: sG holds the (-,+) quadrant of the phase matrix
  double *sY = sG;

  const int t = threadIdx.x;
  const int li = itf / d.nf, f = itf % d.nf;
  const int col = w.item_col[li], z = w.item_layer[li];
  if (t < NMU) {
    smu[t] = w.mu[t];
    sqw[t] = w.qw[t];
  }
  if (w.itf_status[itf] & 0xff) return true;
  const int nleg = w.nleg[itf];
  double kext = w.kext[itf], salb = w.salb[itf];
  if (nleg <= 0) {  // no Legendre series: no scatcnv, zero optics
    kext = 0.0;
    salb = 0.0;
  }
  // ---- scatcnv: the phase matrix, formed by k_phase in the unit's T
  // ((+,+) quadrant) and R ((-,+) quadrant) slots
  const int numpts = 2 * ((nleg + 1) / 2) + 4;
  {
    const bool has = (nleg > 0 || s.tmatrix) && s.lhyd_scat, pf = phase_in_f32(s, w);
    const double *P0 = w.T + (size_t)itf * NN2, *P1 = w.R + (size_t)itf * NN2;
    for (int e = t; e < NN2; e += RT_WG) {
      const int i = e / NN, j = e % NN;
      sX[i * LD + j] = has ? phase_at(pf, P0, e) : 0.0;
      sY[i * LD + j] = has ? phase_at(pf, P1, e) : 0.0;
    }
  }
  __syncthreads();
  const double absorb = s.lhyd_emis ? (1.0 - salb) * kext : 0.0;
  // the T-matrix producer's per-angle optics (tmatrix.f90:640-672) in place
  // of the isotropic kext and (1 - salb) kext: extinction [[k11, k12],
  // [k12, k11]] and emission (e1, e2) of each angle
  const double *tmx = s.tmatrix ? w.tm_ext + (size_t)itf * 2 * NMU : nullptr;
  const double *tme = s.tmatrix ? w.tm_emis + (size_t)itf * 2 * NMU : nullptr;
  // ---- CHECK_NORM4
  if (t < NMU) {
    double sum = tmx ? (s.lhyd_emis ? tme[2 * t] : 0.0) - tmx[2 * t] : absorb - kext;
    for (int j2 = 0; j2 < NMU; j2++)
      sum = sum + 2.0 * PI * sqw[j2] * (sX[(NST * j2) * LD + NST * t] + sY[(NST * j2) * LD + NST * t]);
    snorm[t] = fabs(sum);
  }
  __syncthreads();
  if (t == 0) {
    double mx = 0.0;
    for (int j = 0; j < NMU; j++) mx = fmax(mx, snorm[j]);
    if (mx > 5.0e-4) set_err(&w.itf_status[itf], E_NORM);
  }
  // ---- layer thickness, doubling count (radtran4.f90:311-329)
  const double wl = C_LIGHT / (freqs[f] * 1.e3);
  const double zdiff = fabs(w.hgt_lev[(size_t)col * (d.ML + 1) + z + 1] -
                            w.hgt_lev[(size_t)col * (d.ML + 1) + z]);
  const double gas = fmax(w.kgas[((size_t)col * d.nf + f) * d.ML + z], 0.0);
  const double p0 = planck(w.temp_lev[(size_t)col * (d.ML + 1) + z + 1], wl);
  const double p1 = planck(w.temp_lev[(size_t)col * (d.ML + 1) + z], wl);
  const double extinct = (tmx ? tmx[0] : kext) + gas;
  const double fnd = log(fmax(extinct * zdiff, 1.0e-7) / 1.0e-6) / (double)logf(2.0f);
  const int nd = (fnd > 0.0) ? (int)fnd + 1 : 0;
  const double nsub = (double)ldexpf(1.0f, nd);
  const double dz = zdiff / nsub;
  double linfac = (p0 == 0.0) ? 0.0 : (p1 / p0 - 1.0) / nsub;
  // ---- INITIALIZE4 / INITIAL_SOURCE4
  const double C2 = 2.0 * 3.1415926535897932384;
  for (int e = t; e < NN2; e += RT_WG) {
    const int r = e / NN, c = e % NN;
    const int i2 = r % NST, j2 = r / NST, i1 = c % NST, j1 = c / NST;
    const double tmp = dz / smu[j2];
    const double gext = (i1 == i2) ? gas : 0.0;
    const double diag = (r == c) ? 1.0 : 0.0;
    const double ext =
        (j1 == j2) ? (tmx ? tmx[2 * j2 + (i1 == i2 ? 0 : 1)] : ((i1 == i2) ? kext : 0.0)) + gext : 0.0;
    sR[r * LD + c] = C2 * tmp * sqw[j1] * sY[r * LD + c];
    sT[r * LD + c] = diag - tmp * (ext - C2 * sqw[j1] * sX[r * LD + c]);
  }
  if (t < NN) {
    const int i = t % NST, j = t / NST;
    const double e = (i == 0) ? gas : 0.0;
    const double tmp = p0 * dz / smu[j];
    const double v =
        tmp * ((tmx ? (s.lhyd_emis ? tme[2 * j + i] : 0.0) : (i == 0 ? absorb : 0.0)) + e);
    sv[0][t] = v;  // S+
    sv[1][t] = v;  // S-
    sv[2][t] = v;  // C+
    sv[3][t] = v;  // C-
  }
  __syncthreads();
  // ---- doubling (DOUBLING_INTEGRATION4, SYMMETRIC)
  int nprod = 0;
  for (int it = 0; it < nd; it++) {
    if (!inv_i_minus_ab<GJ, KUNR_DBL>(sR, sR, sX, sG, sinv, nprod)) {
      if (!GJ) return false;
      if (t == 0) set_err(&w.itf_status[itf], E_SINGULAR);
      return true;
    }
    // R' = R + T G R T and T' = T G T with Z = T G and V = R T:
    // R' = R + Z V, T' = Z T (four products, two fused passes).
    mm2<KUNR_DBL>(sT, sG, nullptr, sG, sR, sT, nullptr, sX, nullptr);  // Z = T G (into G), V = R T
    // sources: S+' = (S+ + Z (S+ + R (S- + f C-))) + f C+, C+' = C+ + Z (C+ + R C-),
    // S-' = S- + Z ((S- + f C-) + R S+), C-' = C- + Z (C- + R C+)
    {
      const int g = vl_group(), i = vl_row();
      const bool lead = vl_lead();
      if (t < NN) sv[4][t] = sv[1][t] + linfac * sv[3][t];  // S- + f C-
      __syncthreads();
      const double *v1 = g == 0 ? sv[4] : g == 1 ? sv[3] : g == 2 ? sv[0] : sv[2];
      const double base1 = g == 0 ? sv[0][i] : g == 1 ? sv[2][i] : g == 2 ? sv[4][i] : sv[3][i];
      const double u = base1 + mv_vl(sR, v1, i);
      if (lead) sv[8 + g][i] = u;
      __syncthreads();
      const double z = mv_vl(sG, sv[8 + g], i);
      const double nv = g == 0 ? (sv[0][i] + z) + linfac * sv[2][i]
                               : (g == 1 ? sv[2][i] : g == 2 ? sv[1][i] : sv[3][i]) + z;
      __syncthreads();
      if (lead) sv[g == 0 ? 0 : (g == 1 ? 2 : (g == 2 ? 1 : 3))][i] = nv;
    }
    mm2<KUNR_DBL>(sG, sT, nullptr, sT, sG, sX, sR, sR, nullptr);  // T' = Z T, R' = R + Z V
    nprod += 4;
    linfac = 2.0 * linfac;
  }
  // ---- store R, T, S+, S-
  double *R = w.R + (size_t)itf * NN2, *T = w.T + (size_t)itf * NN2;
  for (int e = t; e < NN2; e += RT_WG) {
    int i = e / NN, j = e % NN;
    R[e] = sR[i * LD + j];
    T[e] = sT[i * LD + j];
  }
  if (t < NN) {
    w.Sp[(size_t)itf * NN + t] = sv[0][t];
    w.Sm[(size_t)itf * NN + t] = sv[1][t];
  }
  if (t == 0) {
    w.ndbl[itf] = nd;
    atomicAdd(&w.counters[3], (unsigned long long)nd);
    if (nleg > 0 && s.lhyd_scat)
      atomicAdd(&w.counters[8], (unsigned long long)(numpts / 2 + 1) * (nleg + 1));
    atomicAdd(&w.counters[9], (unsigned long long)nprod);
    if (GJ) atomicAdd(&w.counters[11], 1ull);
  }
  return true;
}

// ---- the hot kernel's unit on three LDS matrices
//
// The layer doubling needs R, T and a scratch matrix in LDS when R's own
// output tile and V = R T are kept in registers across a step and the inverse
// series is applied to T directly:
//   Z = T G = T (I + X)(I + X^2)(I + X^4)...,  X = R R,
// the factors in the order of series_inv (Z_0 = T + T X, then Z <- Z + Z P,
// P <- P P).  Slot A holds R (then Z, then R'), B holds T, C holds X, the
// powers P and then V.  At 30 KB of LDS five workgroups share a compute unit
// (the four-matrix layout admits four).
struct DblLds3 {
  double A[MSZ], B[MSZ], C[MSZ];
  double v[12][NN];  // S+, S-, C+, C-, then source-chain scratch
  double norm[NMU];
  float wmax[NWAVE];
  double mu[NMU], qw[NMU];
  int slot;  // unit index of a persistent workgroup
};

// Row i of A (x + f y) in the four-vector mapping (y may be null); the lane
// split of mv_pair at N = 32, one lane per row at N = 64.  No barrier.
__device__ __forceinline__ double mv_vl_f(const double *A, const double *x, const double *y,
                                          double f, int i) {
  double s = 0.0;
  if constexpr (VL == 2) {
    const int h = (threadIdx.x & 1) * 8;
    // unrolled by 4 only: the passes around it keep their own operands live
#pragma unroll 4
    for (int k = 0; k < NN / 2; k++) {
      const int c = (k & 7) + h + 16 * (k >> 3);
      s += A[i * LD + c] * (y ? x[c] + f * y[c] : x[c]);
    }
    return s + __shfl_xor(s, 1, 64);
  } else {
#pragma unroll 8
    for (int k = 0; k < NN; k++) s += A[i * LD + k] * (y ? x[k] + f * y[k] : x[k]);
    return s;
  }
}

// One step of DOUBLING_INTEGRATION4 (radintg4.f90:174-309) on the slots of
// DblLds3 (R in A, T in B on entry and on exit).  With V = R T:
//   R' = R + Z V,  T' = Z T,
//   S+' = (S+ + Z (S+ + R (S- + f C-))) + f C+,  C+' = C+ + Z (C+ + R C-),
//   S-' = S- + Z ((S- + f C-) + R S+),           C-' = C- + Z (C- + R C+).
// Returns false when the series has not converged (the unit is queued for
// the Gauss-Jordan pass, nothing of it stored).
template <int U>
__device__ __forceinline__ bool dbl_step3(double *sA, double *sB, double *sC,
                                          double (*sv)[NN], double linfac, float *wmax,
                                          int &np) {
  const int t = threadIdx.x, wv = t >> 6, l = t & 63;
  const int li = l & 15, lk = l >> 4;
  const int g = vl_group(), i = vl_row();
  const bool lead = vl_lead();
  // ---- X = R R into C; V = R T and R's own tile kept in registers; the
  // source chains' first stage beside them
  d4_t ax[TPW], av[TPW], rd[TPW];
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      ax[q][r] = av[q][r] = 0.0;
      rd[q][r] = sA[(I + 4 * r + lk) * LD + J + li];
    }
  }
#pragma unroll U
  for (int s = 0; s < NN / 4; s++) {
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int I = tile_i(wv, q), J = tile_j(wv, q);
      const double a = sA[(I + li) * LD + 4 * s + lk];
      const double bx = sA[(4 * s + lk) * LD + J + li];
      const double bv = sB[(4 * s + lk) * LD + J + li];
      ax[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bx, ax[q], 0, 0, 0);
      av[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, av[q], 0, 0, 0);
    }
  }
  double u;
  {
    // g = 0: S+ + R (S- + f C-); 1: C+ + R C-; 2: (S- + f C-) + R S+; 3: C- + R C+
    const double *x = g == 0 ? sv[1] : g == 1 ? sv[3] : g == 2 ? sv[0] : sv[2];
    const double *y = g == 0 ? sv[3] : nullptr;
    const double b = g == 0 ? sv[0][i] : g == 1 ? sv[2][i] : g == 2 ? sv[1][i] + linfac * sv[3][i]
                                                                    : sv[3][i];
    u = b + mv_vl_f(sA, x, y, linfac, i);
  }
  __syncthreads();
  double m = 0.0;
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) sC[(I + 4 * r + lk) * LD + J + li] = ax[q][r];
    m = max_nan(m, max_nan(max_nan(fabs(ax[q][0]), fabs(ax[q][1])),
                           max_nan(fabs(ax[q][2]), fabs(ax[q][3]))));
  }
  if (lead) sv[8 + g][i] = u;
  {
    const float mf = wave_max((float)m);
    if (l == 0) wmax[wv] = mf;
  }
  __syncthreads();
  m = (double)wmax_all(wmax);
  np += 2;
  if (!(m < 0x1p64)) return false;  // NaN or overflow
  if ((double)NN * m * m < NEU_TOL) {
    // ---- thin layer (G = I + X): Z = T + T X into A, V into C
    d4_t az[TPW];
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
      for (int r = 0; r < 4; r++) az[q][r] = sB[(I + 4 * r + lk) * LD + J + li];
    }
#pragma unroll U
    for (int s = 0; s < NN / 4; s++) {
#pragma unroll
      for (int q = 0; q < TPW; q++) {
        const int I = tile_i(wv, q), J = tile_j(wv, q);
        const double a = sB[(I + li) * LD + 4 * s + lk];
        const double b = sC[(4 * s + lk) * LD + J + li];
        az[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, az[q], 0, 0, 0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int e = (I + 4 * r + lk) * LD + J + li;
        sA[e] = az[q][r];
        sC[e] = av[q][r];
      }
    }
    __syncthreads();
    np += 1;
  } else {
    // ---- thick layer: Z = T + T X and P = X X in one pass, then the
    // remaining factors Z <- Z + Z P, P <- P P; then V into C
    m = mm2<U>(sB, sC, sB, sA, sC, sC, nullptr, sC, wmax);
    np += 2;
    bool conv = false;
    for (int it = 0; it < NEU_MAXIT; it++) {
      if (!(m < 0x1p64)) return false;
      if (m < NEU_TOL) {
        conv = true;
        break;
      }
      if ((double)NN * m * m < NEU_TOL) {  // last factor
        mm_add<U>(sA, sA, sC, sA);
        np++;
        conv = true;
        break;
      }
      m = mm2<U>(sA, sC, sA, sA, sC, sC, nullptr, sC, wmax);
      np += 2;
    }
    if (!conv) return false;
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
      for (int r = 0; r < 4; r++) sC[(I + 4 * r + lk) * LD + J + li] = av[q][r];
    }
    __syncthreads();
  }
  // ---- T' = Z T, R' = R + Z V from the same Z operand loads; the chains'
  // second stage Z u beside them
  d4_t at[TPW];
#pragma unroll
  for (int q = 0; q < TPW; q++)
#pragma unroll
    for (int r = 0; r < 4; r++) at[q][r] = 0.0;
#pragma unroll U
  for (int s = 0; s < NN / 4; s++) {
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int I = tile_i(wv, q), J = tile_j(wv, q);
      const double a = sA[(I + li) * LD + 4 * s + lk];
      const double bt = sB[(4 * s + lk) * LD + J + li];
      const double bv = sC[(4 * s + lk) * LD + J + li];
      at[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bt, at[q], 0, 0, 0);
      rd[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, rd[q], 0, 0, 0);
    }
  }
  double nv;
  {
    const double z = mv_vl_f(sA, sv[8 + g], nullptr, 0.0, i);
    nv = g == 0 ? (sv[0][i] + z) + linfac * sv[2][i]
                : (g == 1 ? sv[2][i] : g == 2 ? sv[1][i] : sv[3][i]) + z;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int e = (I + 4 * r + lk) * LD + J + li;
      sB[e] = at[q][r];
      sA[e] = rd[q][r];
    }
  }
  if (lead) sv[g == 0 ? 0 : (g == 1 ? 2 : (g == 2 ? 1 : 3))][i] = nv;
  __syncthreads();
  np += 2;
  return true;
}

// One (scattering layer, channel) unit of the hot kernel: as doubling_unit
// (scatcnv's phase matrix from k_phase, CHECK_NORM4, the doubling count,
// INITIALIZE4 / INITIAL_SOURCE4, the doublings) on DblLds3.  Returns false
// when an inverse series did not converge; nothing is written for the unit.
__device__ bool doubling_unit3(const Dims &d, const Settings &s, const double *freqs,
                               const Work &w, int itf, DblLds3 &L) {
  double *sA = L.A, *sB = L.B, *sC = L.C;
  double(*sv)[NN] = L.v;
  double *snorm = L.norm, *smu = L.mu, *sqw = L.qw;
  const int t = threadIdx.x;
  const int li = itf / d.nf, f = itf % d.nf;
  const int col = w.item_col[li], z = w.item_layer[li];
  if (t < NMU) {
    smu[t] = w.mu[t];
    sqw[t] = w.qw[t];
  }
  if (w.itf_status[itf] & 0xff) return true;
  const int nleg = w.nleg[itf];
  double kext = w.kext[itf], salb = w.salb[itf];
  if (nleg <= 0) {  // no Legendre series: no scatcnv, zero optics
    kext = 0.0;
    salb = 0.0;
  }
  // ---- scatcnv: the phase matrix, formed by k_phase in the unit's T
  // ((+,+) quadrant, into B) and R ((-,+) quadrant, into A) slots
  const int numpts = 2 * ((nleg + 1) / 2) + 4;
  {
    const bool has = (nleg > 0 || s.tmatrix) && s.lhyd_scat, pf = phase_in_f32(s, w);
    const double *P0 = w.T + (size_t)itf * NN2, *P1 = w.R + (size_t)itf * NN2;
    for (int e = t; e < NN2; e += RT_WG) {
      const int i = e / NN, j = e % NN;
      sB[i * LD + j] = has ? phase_at(pf, P0, e) : 0.0;
      sA[i * LD + j] = has ? phase_at(pf, P1, e) : 0.0;
    }
  }
  __syncthreads();
  const double absorb = s.lhyd_emis ? (1.0 - salb) * kext : 0.0;
  // the T-matrix producer's per-angle optics (tmatrix.f90:640-672) in place
  // of the isotropic kext and (1 - salb) kext: extinction [[k11, k12],
  // [k12, k11]] and emission (e1, e2) of each angle
  const double *tmx = s.tmatrix ? w.tm_ext + (size_t)itf * 2 * NMU : nullptr;
  const double *tme = s.tmatrix ? w.tm_emis + (size_t)itf * 2 * NMU : nullptr;
  // ---- CHECK_NORM4
  if (t < NMU) {
    double sum = tmx ? (s.lhyd_emis ? tme[2 * t] : 0.0) - tmx[2 * t] : absorb - kext;
    for (int j2 = 0; j2 < NMU; j2++)
      sum = sum + 2.0 * PI * sqw[j2] * (sB[(NST * j2) * LD + NST * t] + sA[(NST * j2) * LD + NST * t]);
    snorm[t] = fabs(sum);
  }
  __syncthreads();
  if (t == 0) {
    double mx = 0.0;
    for (int j = 0; j < NMU; j++) mx = fmax(mx, snorm[j]);
    if (mx > 5.0e-4) set_err(&w.itf_status[itf], E_NORM);
  }
  // ---- layer thickness, doubling count (radtran4.f90:311-329)
  const double wl = C_LIGHT / (freqs[f] * 1.e3);
  const double zdiff = fabs(w.hgt_lev[(size_t)col * (d.ML + 1) + z + 1] -
                            w.hgt_lev[(size_t)col * (d.ML + 1) + z]);
  const double gas = fmax(w.kgas[((size_t)col * d.nf + f) * d.ML + z], 0.0);
  const double p0 = planck(w.temp_lev[(size_t)col * (d.ML + 1) + z + 1], wl);
  const double p1 = planck(w.temp_lev[(size_t)col * (d.ML + 1) + z], wl);
  const double extinct = (tmx ? tmx[0] : kext) + gas;
  const double fnd = log(fmax(extinct * zdiff, 1.0e-7) / 1.0e-6) / (double)logf(2.0f);
  const int nd = (fnd > 0.0) ? (int)fnd + 1 : 0;
  const double nsub = (double)ldexpf(1.0f, nd);
  const double dz = zdiff / nsub;
  double linfac = (p0 == 0.0) ? 0.0 : (p1 / p0 - 1.0) / nsub;
  // ---- INITIALIZE4 / INITIAL_SOURCE4, in place: R from the (-,+) quadrant
  // in A, T from the (+,+) quadrant in B
  const double C2 = 2.0 * 3.1415926535897932384;
  for (int e = t; e < NN2; e += RT_WG) {
    const int r = e / NN, c = e % NN;
    const int i2 = r % NST, j2 = r / NST, i1 = c % NST, j1 = c / NST;
    const double tmp = dz / smu[j2];
    const double gext = (i1 == i2) ? gas : 0.0;
    const double diag = (r == c) ? 1.0 : 0.0;
    const double ext =
        (j1 == j2) ? (tmx ? tmx[2 * j2 + (i1 == i2 ? 0 : 1)] : ((i1 == i2) ? kext : 0.0)) + gext : 0.0;
    sA[r * LD + c] = C2 * tmp * sqw[j1] * sA[r * LD + c];
    sB[r * LD + c] = diag - tmp * (ext - C2 * sqw[j1] * sB[r * LD + c]);
  }
  if (t < NN) {
    const int i = t % NST, j = t / NST;
    const double e = (i == 0) ? gas : 0.0;
    const double tmp = p0 * dz / smu[j];
    const double v =
        tmp * ((tmx ? (s.lhyd_emis ? tme[2 * j + i] : 0.0) : (i == 0 ? absorb : 0.0)) + e);
    sv[0][t] = v;  // S+
    sv[1][t] = v;  // S-
    sv[2][t] = v;  // C+
    sv[3][t] = v;  // C-
  }
  __syncthreads();
  // ---- doubling (DOUBLING_INTEGRATION4, SYMMETRIC)
  int nprod = 0;
  for (int it = 0; it < nd; it++) {
    if (!dbl_step3<DBL_STEP_UNR>(sA, sB, sC, sv, linfac, L.wmax, nprod)) return false;
    linfac = 2.0 * linfac;
  }
  // ---- store R, T, S+, S-
  double *R = w.R + (size_t)itf * NN2, *T = w.T + (size_t)itf * NN2;
  for (int e = t; e < NN2; e += RT_WG) {
    int i = e / NN, j = e % NN;
    R[e] = sA[i * LD + j];
    T[e] = sB[i * LD + j];
  }
  if (t < NN) {
    w.Sp[(size_t)itf * NN + t] = sv[0][t];
    w.Sm[(size_t)itf * NN + t] = sv[1][t];
  }
  if (t == 0) {
    w.ndbl[itf] = nd;
    atomicAdd(&w.counters[3], (unsigned long long)nd);
    if (nleg > 0 && s.lhyd_scat)
      atomicAdd(&w.counters[8], (unsigned long long)(numpts / 2 + 1) * (nleg + 1));
    atomicAdd(&w.counters[9], (unsigned long long)nprod);
  }
  return true;
}

// One unit per workgroup, or (PERS) a persistent grid taking units from a
// work queue.  The two are separate instantiations, so that the loop does
// not cost the one-unit kernel registers.
// five waves per SIMD at N = 32 (DblLds3, 28.8 KB: five workgroups per CU);
// at N = 64 its 107 KB admit one workgroup per CU, and the registers are left
// to the larger tiles
constexpr int DBL_WAVES = (NN == 32) ? 5 : 2;
template <bool PERS>
__global__ void __launch_bounds__(RT_WG) __attribute__((amdgpu_waves_per_eu(DBL_WAVES))) k_doubling(Dims d, Settings s, const double *freqs,
                                                    Work w) {
  __shared__ DblLds3 L;
  const int n = d.nitems * d.nf;
  int itf = PERS ? next_unit(&w.queue[1], &L.slot) : (int)blockIdx.x;
  while (itf < n) {
    // inverse_gj == 2 (a test mode) queues every unit for the Gauss-Jordan pass
    const bool done = s.inverse_gj != 2 && doubling_unit3(d, s, freqs, w, itf, L);
    if (!done && threadIdx.x == 0) {
      const unsigned q = atomicAdd(&w.redo_n[0], 1u);
      w.redo_dbl[q] = itf;
    }
    if (!PERS) break;
    itf = next_unit(&w.queue[1], &L.slot);
  }
}

// Gauss-Jordan pass over the queued units (all = 0) or every unit (all = 1)
__global__ void __launch_bounds__(RT_WG) k_doubling_gj(Dims d, Settings s, const double *freqs,
                                                       Work w, int all) {
  __shared__ DblLds L;
  const int n = all ? d.nitems * d.nf : (int)w.redo_n[0];
  if (threadIdx.x == 0) L.inv.gj_only = all;
  for (int q = blockIdx.x; q < n; q += gridDim.x) {
    __syncthreads();
    doubling_unit<true>(d, s, freqs, w, all ? q : w.redo_dbl[q], L);
  }
}


// ------------------------------------------------------------------
// mixed precision: the doubling on the FP32 matrix cores
// ------------------------------------------------------------------
//
// T and G = inv(I - R R) stay close to the identity for thin layers, so they
// are carried as deviations T~ = T - I and G~ = G - I: every entry then has
// single-precision relative accuracy however thin the layer.  With Z~ = T~ +
// (I + T~) G~ and V = R + R T~ a doubling step is
//   T~' = Z~ + (I + Z~) T~,   R' = R + (I + Z~) V,
// four products as in the FP64 kernel; the identity is added to the A
// operand while it is loaded.  The phase matrix, its normalisation check,
// the start layer (a higher-order start, see doubling_unit_f32) and the
// doubling count are computed in FP64; the layer operators are written back
// in FP64 for the adding kernel.

// C = D + (A + iA I) B on the FP32 matrix cores.  16x16x4 f32 layout: lane l
// supplies A[l%16][l/16] and B[l/16][l%16] and holds D[4*(l/16)+r][l%16];
// each wave owns TPW output tiles.  With wmax the return value is max |C_ij|
// (NaN propagates).
template <int U = KUNR_DBL>
__device__ __forceinline__ float mmf(const float *A, bool iA, const float *B, const float *D,
                                     float *C, float *wmax) {
  const int t = threadIdx.x, wv = t >> 6, l = t & 63;
  const int li = l & 15, lk = l >> 4;
  f4_t acc[TPW];
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) acc[q][r] = D ? D[(I + 4 * lk + r) * LD + J + li] : 0.f;
  }
#pragma unroll U
  for (int s = 0; s < NN / 4; s++) {
    const int kc = 4 * s + lk;
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int I = tile_i(wv, q), J = tile_j(wv, q);
      float a = A[(I + li) * LD + kc];
      if (iA && I + li == kc) a += 1.f;
      const float b = B[kc * LD + J + li];
      acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[q], 0, 0, 0);
    }
  }
  __syncthreads();
  float m = 0.f;
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) C[(I + 4 * lk + r) * LD + J + li] = acc[q][r];
    m = max_nan(m, max_nan(max_nan(fabsf(acc[q][0]), fabsf(acc[q][1])),
                           max_nan(fabsf(acc[q][2]), fabsf(acc[q][3]))));
  }
  if (wmax) {
    const float mf = wave_max(m);
    if (l == 0) wmax[wv] = mf;
  }
  __syncthreads();
  return wmax ? wmax_all(wmax) : 0.f;
}

// Two independent FP32 products in one pass (outputs may alias inputs);
// with wmax the return value is max |C2_ij|.
template <int U = KUNR_DBL>
__device__ __forceinline__ float mmf2(const float *A1, bool iA1, const float *B1, const float *D1,
                                      float *C1, const float *A2, bool iA2, const float *B2,
                                      const float *D2, float *C2, float *wmax) {
  const int t = threadIdx.x, wv = t >> 6, l = t & 63;
  const int li = l & 15, lk = l >> 4;
  f4_t acc1[TPW], acc2[TPW];
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      acc1[q][r] = D1 ? D1[(I + 4 * lk + r) * LD + J + li] : 0.f;
      acc2[q][r] = D2 ? D2[(I + 4 * lk + r) * LD + J + li] : 0.f;
    }
  }
#pragma unroll U
  for (int s = 0; s < NN / 4; s++) {
    const int kc = 4 * s + lk;
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int I = tile_i(wv, q), J = tile_j(wv, q);
      const bool dg = I + li == kc;
      float a1 = A1[(I + li) * LD + kc], a2 = A2[(I + li) * LD + kc];
      if (iA1 && dg) a1 += 1.f;
      if (iA2 && dg) a2 += 1.f;
      const float b1 = B1[kc * LD + J + li], b2 = B2[kc * LD + J + li];
      acc1[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc1[q], 0, 0, 0);
      acc2[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, b2, acc2[q], 0, 0, 0);
    }
  }
  __syncthreads();
  float m = 0.f;
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int I = tile_i(wv, q), J = tile_j(wv, q);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      C1[(I + 4 * lk + r) * LD + J + li] = acc1[q][r];
      C2[(I + 4 * lk + r) * LD + J + li] = acc2[q][r];
    }
    m = max_nan(m, max_nan(max_nan(fabsf(acc2[q][0]), fabsf(acc2[q][1])),
                           max_nan(fabsf(acc2[q][2]), fabsf(acc2[q][3]))));
  }
  if (wmax) {
    const float mf = wave_max(m);
    if (l == 0) wmax[wv] = mf;
  }
  __syncthreads();
  return wmax ? wmax_all(wmax) : 0.f;
}

// Row i of A x in FP32 over the lane pair t, t^1 (as mv_pair).
__device__ __forceinline__ float mvf_pair(const float *A, const float *x, int i) {
  const int h = (threadIdx.x & 1) * 8;
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < NN / 2; k++) {
    const int c = (k & 7) + h + 16 * (k >> 3);
    s += A[i * LD + c] * x[c];
  }
  return s + __shfl_xor(s, 1, 64);
}

// Row i of A x in FP32 in the four-vector mapping (as mv_vl)
__device__ __forceinline__ float mvf_vl(const float *A, const float *x, int i) {
  if constexpr (VL == 2) {
    return mvf_pair(A, x, i);
  } else {
    float s = 0.f;
#pragma unroll 8
    for (int k = 0; k < NN; k++) s += A[i * LD + k] * x[k];
    return s;
  }
}

// Truncation of the FP32 series: the omitted tail is below the single-
// precision resolution of G = I + G~.
constexpr float NEU_TOL_F = 0x1p-28f;

// G~ = inv(I - R R) - I by the product-form series in FP32 (X scratch).
// false: not converged (the unit goes to the FP64 Gauss-Jordan pass).
__device__ bool series_f32(const float *R, float *X, float *Gt, float *wmax, int &np) {
  const int t = threadIdx.x;
  np++;
  float m = mmf(R, false, R, nullptr, X, wmax);
  for (int e = t; e < NN2; e += RT_WG) {
    const int i = e / NN, j = e % NN;
    Gt[i * LD + j] = X[i * LD + j];
  }
  __syncthreads();
  if (!(m < 0x1p64f)) return false;
  if ((float)NN * m * m < NEU_TOL_F) return true;
  m = mmf(X, false, X, nullptr, X, wmax);
  np++;
  for (int it = 0; it < NEU_MAXIT; it++) {
    if (!(m < 0x1p64f)) return false;
    if (m < NEU_TOL_F) return true;
    if ((float)NN * m * m < NEU_TOL_F) {  // last factor: G~ <- G~ + (I + G~) P
      mmf(Gt, true, X, Gt, Gt, nullptr);
      np++;
      return true;
    }
    m = mmf2(Gt, true, X, Gt, Gt, X, false, X, nullptr, X, wmax);
    np += 2;
  }
  return false;
}

// Start layer of the mixed-precision path.
//
// The reference doubles from a first-order layer of optical depth <= 1e-6
// (radtran4.f90:323-345, INITIALIZE4 radintg4.f90:1-45), so ~16 doublings
// per layer, and its result carries that start's truncation error (~1e-5
// relative in R).  Here the start layer is the exact solution of the same
// discrete-ordinate system over a sub-layer zdiff / 2^nd, with nd the
// smallest count that makes ||H dz||_inf <= HS_THETA; that removes about 13
// doublings.  It is formed in FP32, like the doublings that follow: its
// truncation error (below) is at the FP32 resolution.
//
// With the first-order operators R1 = dz A, T1 = I - dz B (A, B per unit
// length, from the phase matrix and the extinction) the intensities of the
// two hemispheres obey d(I+, I-)/dt = H (I+, I-) + (e, -e) b(t),
// H = [[-B, A], [-A, B]].  In p = I+ + I-, m = I+ - I- the generator is
// [[0, -S], [-D, 0]] with S = A + B, D = B - A, whose exponential over dz
// has the blocks c(X), s(X) S, s(Y) D, c(Y) with X = S D dz^2,
// Y = D S dz^2, c(Z) = sum Z^k / (2k)!, s(Z) = sum Z^k / (2k+1)!.  The
// transmission and reflection of the sub-layer are then
//   U = c(X) + dz s(X) S,  V = c(Y) + dz s(Y) D,
//   T = ((U + V) / 2)^-1,  R = ((U - V) / 2) T.
// For the Planck function b0 (1 + f t / dz) the particular solution is
// g = sum_n H^n (e, -e) dz^(n+1) [1/(n+1)! + f/(n+2)!], and the sources are
// S+ = g+ - R g-, S- = -T g-.  Doubling from there with the reference's
// linear-source recursion (linfac = f) gives the layer's R, T, S.
// c and s are truncated after Z^2 and the particular sum after n = 7: the
// start's relative error is ~theta^6/720 (3e-7).
constexpr double HS_THETA = 0.25;

struct MixLds {
  // the start's matrices (dz S, dz D, X, Y and their sums), then the
  // doubling's R, T~, G~ and scratch
  float m[4][MSZ];
  float v[16][NN];  // S+, S-, C+, C-, then source-chain scratch
  double g[4][NN];  // particular sums: constant p, m; ramp p, m
  double c[2][NN];  // chain vectors, then g- of the two sources
  double hrow[NN];
  double norm[NMU];
  float wmax[NWAVE];
  double mu[NMU], qw[NMU];
  int slot;  // unit index of a persistent workgroup
};

// Row i of A x for an FP32 matrix and an FP64 vector, accumulated in FP64
// over the lane pair t, t^1 (as mv_pair).
__device__ __forceinline__ double mvfd_pair(const float *A, const double *x, int i) {
  const int h = (threadIdx.x & 1) * 8;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < NN / 2; k++) {
    const int c = (k & 7) + h + 16 * (k >> 3);
    s += (double)A[i * LD + c] * x[c];
  }
  return s + __shfl_xor(s, 1, 64);
}
static_assert(VL == 2, "the pair mapping of the mixed start");

// G~ = inv(I - X) - I by the product-form series in FP32, X given (and
// overwritten by its powers), truncated when max |X^(2^k)| < tol.  false:
// not converged, or a NaN.
__device__ bool series_inv_f32(float *X, float *Gt, float *wmax, float tol, int &np) {
  const int t = threadIdx.x;
  for (int e = t; e < NN2; e += RT_WG) {
    const int i = e / NN, j = e % NN;
    Gt[i * LD + j] = X[i * LD + j];
  }
  float m = mmf(X, false, X, nullptr, X, wmax);  // its barriers order the Gt writes
  np++;
  for (int it = 0; it < NEU_MAXIT; it++) {
    if (!(m < 0x1p64f)) return false;
    if (m < tol) return true;
    if ((float)NN * m * m < tol) {  // last factor: G~ <- G~ + (I + G~) X
      mmf(Gt, true, X, Gt, Gt, nullptr);
      np++;
      return true;
    }
    m = mmf2(Gt, true, X, Gt, Gt, X, false, X, nullptr, X, wmax);
    np += 2;
  }
  return false;
}

// One (scattering layer, channel) unit of the mixed-precision path; false
// when an inverse series did not converge (nothing written; the unit is
// recomputed in FP64 by the Gauss-Jordan pass).  The start layer is formed
// in FP32 like the doublings: its operators are I plus terms of order
// theta, and they are carried as those terms (c(X) - I, U - I, V - I), so
// no FP32 difference cancels an identity.  CHECK_NORM4 and the choice of
// nd read the FP64 phase matrix itself.
__device__ bool doubling_unit_f32(const Dims &d, const Settings &s, const double *freqs,
                                  const Work &w, int itf, MixLds &L) {
  float *M0 = L.m[0], *M1 = L.m[1], *M2 = L.m[2], *M3 = L.m[3];
  double *smu = L.mu, *sqw = L.qw, *snorm = L.norm;

  const int t = threadIdx.x;
  const int li = itf / d.nf, f = itf % d.nf;
  const int col = w.item_col[li], z = w.item_layer[li];
  if (t < NMU) {
    smu[t] = w.mu[t];
    sqw[t] = w.qw[t];
  }
  if (w.itf_status[itf] & 0xff) return true;
  const int nleg = w.nleg[itf];
  double kext = w.kext[itf], salb = w.salb[itf];
  if (nleg <= 0) {
    kext = 0.0;
    salb = 0.0;
  }
  // the phase matrix, formed by k_phase in the unit's T and R slots:
  // P0 = (+,+) quadrant, P1 = (-,+) quadrant
  // staged in FP64 over the LDS of the four FP32 matrices
  const int numpts = 2 * ((nleg + 1) / 2) + 4;
  {
    const bool has = (nleg > 0 || s.tmatrix) && s.lhyd_scat, pf = phase_in_f32(s, w);
    const double *G0 = w.T + (size_t)itf * NN2, *G1 = w.R + (size_t)itf * NN2;
    double *D0 = reinterpret_cast<double *>(L.m[0]), *D1 = reinterpret_cast<double *>(L.m[2]);
    for (int e = t; e < NN2; e += RT_WG) {
      D0[e] = has ? phase_at(pf, G0, e) : 0.0;
      D1[e] = has ? phase_at(pf, G1, e) : 0.0;
    }
  }
  const double *P0 = reinterpret_cast<const double *>(L.m[0]);
  const double *P1 = reinterpret_cast<const double *>(L.m[2]);
  __syncthreads();
  const double absorb = s.lhyd_emis ? (1.0 - salb) * kext : 0.0;
  // the T-matrix producer's per-angle optics (tmatrix.f90:640-672) in place
  // of the isotropic kext and (1 - salb) kext: extinction [[k11, k12],
  // [k12, k11]] and emission (e1, e2) of each angle
  const double *tmx = s.tmatrix ? w.tm_ext + (size_t)itf * 2 * NMU : nullptr;
  const double *tme = s.tmatrix ? w.tm_emis + (size_t)itf * 2 * NMU : nullptr;
  const double wl = C_LIGHT / (freqs[f] * 1.e3);
  const double zdiff = fabs(w.hgt_lev[(size_t)col * (d.ML + 1) + z + 1] -
                            w.hgt_lev[(size_t)col * (d.ML + 1) + z]);
  const double gas = fmax(w.kgas[((size_t)col * d.nf + f) * d.ML + z], 0.0);
  const double p0 = planck(w.temp_lev[(size_t)col * (d.ML + 1) + z + 1], wl);
  const double p1 = planck(w.temp_lev[(size_t)col * (d.ML + 1) + z], wl);
  const double ext = kext + gas;
  const double C2 = 2.0 * 3.1415926535897932384;
  if (t < NMU) {  // CHECK_NORM4
    double sum = tmx ? (s.lhyd_emis ? tme[2 * t] : 0.0) - tmx[2 * t] : absorb - kext;
    for (int j2 = 0; j2 < NMU; j2++) {
      const int e = (NST * j2) * NN + NST * t;
      sum = sum + 2.0 * PI * sqw[j2] * (P0[e] + P1[e]);
    }
    snorm[t] = fabs(sum);
  } else if (t >= 64 && t < 64 + NN) {  // row r of |A| + |B|, per unit length
    const int r = t - 64;
    double sum = 0.0;
    for (int c = 0; c < NN; c++) {
      const double wc = C2 * sqw[c / NST];
      sum += fabs(wc * P1[r * NN + c]) + fabs((c == r ? ext : 0.0) - wc * P0[r * NN + c]);
    }
    if (tmx) {
      // a T-matrix unit's per-angle extinction matrix: its diagonal entry and
      // the same-angle Stokes partner (column r ^ 1) replace ext and 0
      const int j = r / NST, c1 = r ^ 1;
      const double w0 = C2 * sqw[j];
      sum += fabs((tmx[2 * j] + gas) - w0 * P0[r * NN + r]) - fabs(ext - w0 * P0[r * NN + r]);
      sum += fabs(tmx[2 * j + 1] - w0 * P0[r * NN + c1]) - fabs(0.0 - w0 * P0[r * NN + c1]);
    }
    L.hrow[r] = sum / smu[r / NST];
  }
  __syncthreads();
  if (t == 0) {
    double mx = 0.0;
    for (int j = 0; j < NMU; j++) mx = fmax(mx, snorm[j]);
    if (mx > 5.0e-4) set_err(&w.itf_status[itf], E_NORM);
  }
  double hn = 0.0;
  for (int r = 0; r < NN; r++) hn = max_nan(hn, L.hrow[r]);
  int nd = 0;
  while (nd < 64 && hn * zdiff > ldexp(HS_THETA, nd)) nd++;
  const double dz = ldexp(zdiff, -nd);
  double linfac = (p0 == 0.0) ? 0.0 : ldexp(p1 / p0 - 1.0, -nd);
  // dz S into M0, dz D into M1 (formed in FP64, rounded once, written
  // over the staged phase matrix); dz e into the chain vector
  {
    float vs[NN2 / RT_WG], vd[NN2 / RT_WG];
#pragma unroll
    for (int q = 0; q < NN2 / RT_WG; q++) {
      const int e = t + q * RT_WG, r = e / NN, c = e % NN;
      const double sc = dz / smu[r / NST], wc = C2 * sqw[c / NST];
      const double a = sc * wc * P1[e];
      const double b = sc * ((c == r ? ext : 0.0) - wc * P0[e]);
      vs[q] = (float)(b + a);
      vd[q] = (float)(b - a);
    }
    // a T-matrix unit: the two entries of row t that its per-angle extinction
    // matrix changes (diagonal, same-angle Stokes partner), patched below
    float ts[2] = {0.0f, 0.0f}, td[2] = {0.0f, 0.0f};
    if (tmx && t < NN) {
      const int j = t / NST;
      const double sc = dz / smu[j], wc = C2 * sqw[j];
      for (int k = 0; k < 2; k++) {
        const int c = k == 0 ? t : (t ^ 1), e = t * NN + c;
        const double a = sc * wc * P1[e];
        const double b = sc * ((k == 0 ? tmx[2 * j] + gas : tmx[2 * j + 1]) - wc * P0[e]);
        ts[k] = (float)(b + a);
        td[k] = (float)(b - a);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NN2 / RT_WG; q++) {
      const int e = t + q * RT_WG, r = e / NN, c = e % NN;
      M0[r * LD + c] = vs[q];
      M1[r * LD + c] = vd[q];
    }
    if (tmx) {
      __syncthreads();
      if (t < NN) {
        M0[t * LD + t] = ts[0];
        M1[t * LD + t] = td[0];
        M0[t * LD + (t ^ 1)] = ts[1];
        M1[t * LD + (t ^ 1)] = td[1];
      }
    }
  }
  double *cv = L.c[0], *cw = L.c[1];
  if (t < NN)
    cv[t] = tmx ? dz * ((t % NST == 0 ? gas : 0.0) + (s.lhyd_emis ? tme[t] : 0.0)) / smu[t / NST]
                : (t % NST == 0) ? dz * (absorb + gas) / smu[t / NST] : 0.0;
  __syncthreads();
  // particular sums with v_0 = dz e, w_k = dz S v_k, v_k+1 = dz D w_k:
  //   constant: m = 2 sum v_k / (2k+1)!,  p = -2 sum w_k / (2k+2)!
  //   ramp:     m = 2 sum v_k / (2k+2)!,  p = -2 sum w_k / (2k+3)!
  {
    constexpr double ifact[10] = {1.0,         1.0,          0.5,           1.0 / 6.0,
                                  1.0 / 24.0,  1.0 / 120.0,  1.0 / 720.0,   1.0 / 5040.0,
                                  1.0 / 40320.0, 1.0 / 362880.0};
    double gcp = 0.0, gcm = 0.0, grp = 0.0, grm = 0.0;
    const int i = vl_row();
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (t < RT_WG / 4) {  // the first vector group: one chain
        const double v = cv[i];
        gcm += 2.0 * v * ifact[2 * k + 1];
        grm += 2.0 * v * ifact[2 * k + 2];
        const double wk = mvfd_pair(M0, cv, i);
        gcp -= 2.0 * wk * ifact[2 * k + 2];
        grp -= 2.0 * wk * ifact[2 * k + 3];
        if (vl_lead()) cw[i] = wk;
      }
      __syncthreads();
      if (k < 3) {
        if (t < RT_WG / 4) {
          const double v = mvfd_pair(M1, cw, i);
          if (vl_lead()) cv[i] = v;
        }
        __syncthreads();
      }
    }
    if (t < RT_WG / 4 && vl_lead()) {
      L.g[0][i] = gcp;
      L.g[1][i] = gcm;
      L.g[2][i] = grp;
      L.g[3][i] = grm;
    }
  }
  // (the next product's barriers order these writes)
  int nprod = 7;
  mmf(M0, false, M1, nullptr, M2, nullptr);  // X
  mmf(M2, false, M2, nullptr, M3, nullptr);  // X^2
  for (int e = t; e < NN2; e += RT_WG) {
    const int i = e / NN, j = e % NN;
    const float x = M2[i * LD + j], x2 = M3[i * LD + j];
    M2[i * LD + j] = (i == j ? 1.f : 0.f) + (x * (1.f / 6.f) + x2 * (1.f / 120.f));  // s(X)
    M3[i * LD + j] = x * 0.5f + x2 * (1.f / 24.f);                                  // c(X) - I
  }
  __syncthreads();
  mmf(M2, false, M0, M3, M3, nullptr);  // U - I = c(X) - I + s(X) dz S
  mmf(M1, false, M0, nullptr, M2, nullptr);  // Y
  mmf(M2, false, M2, nullptr, M0, nullptr);  // Y^2
  for (int e = t; e < NN2; e += RT_WG) {
    const int i = e / NN, j = e % NN;
    const float y = M2[i * LD + j], y2 = M0[i * LD + j];
    M2[i * LD + j] = (i == j ? 1.f : 0.f) + (y * (1.f / 6.f) + y2 * (1.f / 120.f));  // s(Y)
    M0[i * LD + j] = y * 0.5f + y2 * (1.f / 24.f);                                  // c(Y) - I
  }
  __syncthreads();
  mmf(M2, false, M1, M0, M0, nullptr);  // V - I = c(Y) - I + s(Y) dz D
  for (int e = t; e < NN2; e += RT_WG) {
    const int i = e / NN, j = e % NN;
    const float u = M3[i * LD + j], v = M0[i * LD + j];
    M1[i * LD + j] = -0.5f * (u + v);  // I - (U + V) / 2
    M3[i * LD + j] = 0.5f * (u - v);   // (U - V) / 2
  }
  if (t < NN) {  // g- of the linear and of the constant source
    const double gcmin = 0.5 * (L.g[0][t] - L.g[1][t]), grmin = 0.5 * (L.g[2][t] - L.g[3][t]);
    cv[t] = p0 * (gcmin + linfac * grmin);
    cw[t] = p0 * gcmin;
  }
  __syncthreads();
  // T = I + G~ with G~ = inv(I - M1) - I into M2; R = M3 T into M0
  if (!series_inv_f32(M1, M2, L.wmax, NEU_TOL_F, nprod)) return false;
  mmf(M3, false, M2, M3, M0, nullptr);
  nprod++;
  // S+ = g+ - R g-, S- = -T g-: linear source (groups 0, 1), constant (2, 3)
  float *sR = M0, *sT = M2, *sG = M1, *sX = M3;
  float(*sv)[NN] = L.v;
  {
    const int g = vl_group(), i = vl_row();
    const double *gm = g < 2 ? cv : cw;
    const double mv = (g & 1) ? gm[i] + mvfd_pair(sT, gm, i) : mvfd_pair(sR, gm, i);
    const double gcplus = 0.5 * (L.g[0][i] + L.g[1][i]), grplus = 0.5 * (L.g[2][i] + L.g[3][i]);
    const double gplus = p0 * (g < 2 ? gcplus + linfac * grplus : gcplus);
    const float src = (float)((g & 1) ? -mv : gplus - mv);
    if (vl_lead()) sv[g][i] = src;
  }
  __syncthreads();
  for (int it = 0; it < nd; it++) {
    if (!series_f32(sR, sX, sG, L.wmax, nprod)) return false;
    // Z~ = T~ + (I + T~) G~ (into G), V = R + R T~
    mmf2(sT, true, sG, sT, sG, sR, false, sT, sR, sX, nullptr);
    {
      const float lf = (float)linfac;
      const int g = vl_group(), i = vl_row();
      const bool lead = vl_lead();
      if (t < NN) sv[4][t] = sv[1][t] + lf * sv[3][t];
      __syncthreads();
      const float *v1 = g == 0 ? sv[4] : g == 1 ? sv[3] : g == 2 ? sv[0] : sv[2];
      const float base1 = g == 0 ? sv[0][i] : g == 1 ? sv[2][i] : g == 2 ? sv[4][i] : sv[3][i];
      const float u = base1 + mvf_vl(sR, v1, i);
      if (lead) sv[8 + g][i] = u;
      __syncthreads();
      const float zu = u + mvf_vl(sG, sv[8 + g], i);  // Z u = u + Z~ u
      const float nv = g == 0 ? (sv[0][i] + zu) + lf * sv[2][i]
                              : (g == 1 ? sv[2][i] : g == 2 ? sv[1][i] : sv[3][i]) + zu;
      __syncthreads();
      if (lead) sv[g == 0 ? 0 : (g == 1 ? 2 : (g == 2 ? 1 : 3))][i] = nv;
    }
    // T~' = Z~ + (I + Z~) T~, R' = R + (I + Z~) V
    mmf2(sG, true, sT, sG, sT, sG, true, sX, sR, sR, nullptr);
    nprod += 4;
    linfac = 2.0 * linfac;
  }
  double *R = w.R + (size_t)itf * NN2, *T = w.T + (size_t)itf * NN2;
  for (int e = t; e < NN2; e += RT_WG) {
    int i = e / NN, j = e % NN;
    R[e] = (double)sR[i * LD + j];
    T[e] = (i == j ? 1.0 : 0.0) + (double)sT[i * LD + j];
  }
  if (t < NN) {
    w.Sp[(size_t)itf * NN + t] = (double)sv[0][t];
    w.Sm[(size_t)itf * NN + t] = (double)sv[1][t];
  }
  if (t == 0) {
    w.ndbl[itf] = nd;
    atomicAdd(&w.counters[3], (unsigned long long)nd);
    if (nleg > 0 && s.lhyd_scat)
      atomicAdd(&w.counters[8], (unsigned long long)(numpts / 2 + 1) * (nleg + 1));
    atomicAdd(&w.counters[9], (unsigned long long)nprod);
  }
  return true;
}

// At N = 32 one workgroup per unit fits seven per compute unit in LDS
// (21 KB each); the register budget follows.
template <bool PERS>
__global__ void __launch_bounds__(RT_WG) __attribute__((amdgpu_waves_per_eu((NN == 32 && !PERS) ? 7 : 1))) k_doubling_mixed(Dims d, Settings s, const double *freqs,
                                                          Work w) {
  __shared__ MixLds L;
  const int n = d.nitems * d.nf;
  int itf = PERS ? next_unit(&w.queue[1], &L.slot) : (int)blockIdx.x;
  while (itf < n) {
    if (!doubling_unit_f32(d, s, freqs, w, itf, L) && threadIdx.x == 0) {
      const unsigned q = atomicAdd(&w.redo_n[0], 1u);
      w.redo_dbl[q] = itf;
    }
    if (!PERS) break;
    itf = next_unit(&w.queue[1], &L.slot);
  }
}

// ------------------------------------------------------------------
// adding, internal radiance, TB
// ------------------------------------------------------------------

// operator of the layer stack above the current level: R+, R-, T+, T- (LDS
// matrices) and S+, S- (LDS vectors)
template <typename T>
struct StackT {
  T *Rp, *Rm, *Tp, *Tm, *Sp, *Sm;
};
typedef StackT<double> Stack;

// Row i = t/RPL of A x on all 256 threads: lane t%RPL sums the NN/RPL
// columns from (NN/RPL)(t%RPL) on (4 at N = 32, conflict-free LDS banks; 16
// at N = 64), the RPL lanes of a row combine by shuffles; all of them return
// the full sum.  No barrier.
constexpr int MVC = NN / RPL;
template <typename T>
__device__ __forceinline__ T mv_row8(const T *A, const T *x, int i) {
  const int c0 = MVC * (threadIdx.x % RPL);
  T s = A[i * LD + c0] * x[c0];
#pragma unroll
  for (int k = 1; k < MVC; k++) s += A[i * LD + c0 + k] * x[c0 + k];
#pragma unroll
  for (int o = 1; o < RPL; o <<= 1) s += __shfl_xor(s, o, 64);
  return s;
}
__device__ __forceinline__ int r8_row() { return threadIdx.x / RPL; }
__device__ __forceinline__ bool r8_lead() { return (threadIdx.x % RPL) == 0; }

// v <- A v on all threads, with barriers so that the shared vector may be
// reused between stages
template <typename T>
__device__ __forceinline__ void mv_stage(const T *A, T *v) {
  const T u = mv_row8(A, v, r8_row());
  __syncthreads();
  if (r8_lead()) v[r8_row()] = u;
  __syncthreads();
}

// COMBINE_LAYERS4 (radintg4.f90:447-513): stack on top of a doubled layer
// with R2+ = R2- = R2 and T2+ = T2- = T2.  The minus side goes first, since
// it reads the stack's old R+ and T+.  Each side is grouped as
//   A = T1m G,  R1m' = R1m + A (R2 T1p),  T1m' = A T2,  S1m' = S1m + A (...)
//   B = T2 G',  R1p' = R2 + B (R1p T2),   T1p' = B T1p,  S1p' = S2p + B (...)
// four products in two fused passes per side (the reference's order of the
// products differs only by association).  T = float: the mixed path, every
// operand and product in FP32.
template <bool GJ, typename T>
__device__ bool combine_scat(StackT<T> &s1, const T *R2, const T *T2, const T *S2p, const T *S2m,
                             T *G, T *X, T *vs, InvLds &sinv, int &np) {
  const int t = threadIdx.x;
  if (!inv_i_minus_ab<GJ>(R2, s1.Rp, X, G, sinv, np)) return false;
  {
    const T u = mv_row8(R2, s1.Sp, r8_row());
    if (r8_lead()) vs[r8_row()] = S2m[r8_row()] + u;  // S2m + R2m S1p
  }
  mm2(s1.Tm, G, nullptr, X, R2, s1.Tp, nullptr, G, nullptr);  // X = A, G = R2 T1p
  mv_stage(X, vs);
  if (t < NN) s1.Sm[t] = s1.Sm[t] + vs[t];                     // S1m + A (...)
  mm2(X, G, s1.Rm, s1.Rm, X, T2, nullptr, s1.Tm, nullptr);  // R1m', T1m'
  np += 4;
  if (!inv_i_minus_ab<GJ>(s1.Rp, R2, X, G, sinv, np)) return false;
  {
    const T u = mv_row8(s1.Rp, S2m, r8_row());
    if (r8_lead()) vs[r8_row()] = s1.Sp[r8_row()] + u;  // S1p + R1p S2m
  }
  mm2(T2, G, nullptr, X, s1.Rp, T2, nullptr, G, nullptr);  // X = B, G = R1p T2
  mv_stage(X, vs);
  if (t < NN) s1.Sp[t] = S2p[t] + vs[t];                       // S2p + B (...)
  mm2(X, G, R2, s1.Rp, X, s1.Tp, nullptr, s1.Tp, nullptr);  // R1p', T1p'
  np += 4;
  __syncthreads();
  return true;
}

// The same combine for a clear layer (NONSCATTER_LAYER4: R2 = 0,
// T2 = diag(dg)).  With R2 = 0 both gammas are the identity and every
// product reduces to a diagonal scaling or one matrix-vector product.
template <typename T>
__device__ void combine_clear(StackT<T> &s1, const T *dg, const T *S2p, const T *S2m) {
  const int t = threadIdx.x, i8 = r8_row();
  const T u = mv_row8(s1.Tm, S2m, i8);                // T1m S2m
  const T u2 = s1.Sp[i8] + mv_row8(s1.Rp, S2m, i8);  // S1p + R1p S2m
  __syncthreads();
  if (r8_lead()) {
    s1.Sm[i8] = s1.Sm[i8] + u;
    s1.Sp[i8] = S2p[i8] + dg[i8] * u2;
  }
  for (int e = t; e < NN2; e += RT_WG) {
    const int i = e / NN, j = e % NN;
    s1.Rp[i * LD + j] = dg[i] * (s1.Rp[i * LD + j] * dg[j]);
    s1.Tp[i * LD + j] = dg[i] * s1.Tp[i * LD + j];
    s1.Tm[i * LD + j] = s1.Tm[i * LD + j] * dg[j];
  }
  __syncthreads();
}

// The combine with the surface "layer" of get_sfc_matrices
// (sfc_matrices.f90:21-194): R2+ = 0, R2- = Rs, T2 = I, S2 = 0.
template <bool GJ, typename T>
__device__ bool combine_surface(StackT<T> &s1, const T *Rs, T *G, T *X, T *vs, InvLds &sinv,
                                int &np) {
  const int t = threadIdx.x;
  if (!inv_i_minus_ab<GJ>(Rs, s1.Rp, X, G, sinv, np)) return false;
  {
    const T u = mv_row8(Rs, s1.Sp, r8_row());
    if (r8_lead()) vs[r8_row()] = u;
  }
  __syncthreads();
  mv_stage(G, vs);
  mv_stage(s1.Tm, vs);
  if (t < NN) s1.Sm[t] = s1.Sm[t] + vs[t];
  mm(Rs, s1.Tp, X);
  mm(G, X, X);
  mm_add(s1.Rm, s1.Tm, X, s1.Rm);
  mm(s1.Tm, G, s1.Tm);
  if (!inv_i_minus_ab<GJ>(s1.Rp, Rs, X, G, sinv, np)) return false;
  if (t < NN) vs[t] = s1.Sp[t];
  __syncthreads();
  mv_stage(G, vs);
  if (t < NN) s1.Sp[t] = vs[t];
  mm(G, s1.Rp, s1.Rp);
  mm(G, s1.Tp, s1.Tp);
  np += 6;
  return true;
}

// The next layer's inputs, fetched into registers while the current layer
// is combined (the barriers of the combine do not wait for global loads).
// At N = 32 the R and T elements come ahead into registers (8 per thread);
// at N = 64 (16 per thread of 512, in kernels already at the register limit)
// stage_layer copies them from HBM itself.
constexpr bool FETCH_RT = (NN == 32);
constexpr int FETCH_N = FETCH_RT ? NN2 / RT_WG : 1;
struct LayerFetch {
  int itf;                       // scattering unit, or -1 for a clear layer
  int st;                        // its status (thread 0)
  double r[FETCH_N], tt[FETCH_N];  // R, T elements t + RT_WG k (FETCH_RT)
  const double *gR, *gT;         // the unit's R and T in HBM (N = 64)
  double sp, sm;                 // S+, S- (t < NN)
  double h0, h1, kg, tp0, tp1;   // clear layer: levels, gas, temperatures
};

__device__ __forceinline__ void fetch_layer(const Dims &d, const Work &w, int col, int f, int z,
                                            LayerFetch &F) {
  const int t = threadIdx.x;
  const int item = w.layer_item[(size_t)col * d.ML + z];
  if (item >= 0) {
    const size_t itf = (size_t)(item - d.item0) * d.nf + f;
    F.itf = (int)itf;
    const double *R = w.R + itf * NN2, *T = w.T + itf * NN2;
    F.gR = R;
    F.gT = T;
    if constexpr (FETCH_RT) {
#pragma unroll
      for (int k = 0; k < NN2 / RT_WG; k++) {
        F.r[k] = R[t + RT_WG * k];
        F.tt[k] = T[t + RT_WG * k];
      }
    }
    if (t < NN) {
      F.sp = w.Sp[itf * NN + t];
      F.sm = w.Sm[itf * NN + t];
    }
    if (t == 0) F.st = w.itf_status[itf];
  } else {
    F.itf = -1;
    const size_t ol = (size_t)col * (d.ML + 1) + z;
    F.h0 = w.hgt_lev[ol];
    F.h1 = w.hgt_lev[ol + 1];
    F.kg = w.kgas[((size_t)col * d.nf + f) * d.ML + z];
    F.tp0 = w.temp_lev[ol];
    F.tp1 = w.temp_lev[ol + 1];
  }
}

// Stages a fetched layer: a doubled layer into (R2, T2, S2p, S2m), or the
// closed form of a clear layer into (dg, S2p, S2m).  Returns true when the
// layer scatters.
template <typename T>
__device__ bool stage_layer(const LayerFetch &F, double wl, const double *smu, T *R2, T *T2, T *dg,
                            T *S2p, T *S2m, int *sbad) {
  const int t = threadIdx.x;
  if (F.itf >= 0) {
    if (t == 0 && (F.st & 0xff)) atomicMax(sbad, F.st);
    if constexpr (FETCH_RT) {
#pragma unroll
      for (int k = 0; k < NN2 / RT_WG; k++) {
        const int e = t + RT_WG * k, i = e / NN, j = e % NN;
        R2[i * LD + j] = F.r[k];
        T2[i * LD + j] = F.tt[k];
      }
    } else {
#pragma unroll 4
      for (int k = 0; k < NN2 / RT_WG; k++) {
        const int e = t + RT_WG * k, i = e / NN, j = e % NN;
        R2[i * LD + j] = F.gR[e];
        T2[i * LD + j] = F.gT[e];
      }
    }
    if (t < NN) {
      S2p[t] = F.sp;
      S2m[t] = F.sm;
    }
    __syncthreads();
    return true;
  }
  // NONSCATTER_LAYER4 (radintg4.f90:79-117)
  const double zdiff = fabs(F.h1 - F.h0);
  const double gas = fmax(F.kg, 0.0);
  const double dtau = zdiff * gas;
  const double p0 = planck(F.tp1, wl);  // top of the layer
  const double p1 = planck(F.tp0, wl);  // bottom
  if (t < NN) {
    const int i = t % NST, j = t / NST;
    dg[t] = exp(-dtau / smu[j]);
    double a = 0.0, b = 0.0;
    if (i == 0 && dtau > 0.0) {
      const double path = dtau / smu[j];
      const double slope = (p1 - p0) / path;
      a = p1 - slope - (p1 - slope * (1.0 + path)) * exp(-path);
      b = p0 + slope - (p0 + slope * (1.0 + path)) * exp(-path);
    }
    S2p[t] = a;
    S2m[t] = b;
  }
  __syncthreads();
  return false;
}

// INTERNAL_RADIANCE4 at the surface level: up = stack of all layers, down =
// the surface (R- = Rs, T- = I, S- = 0)
template <bool GJ, typename T>
__device__ bool radiance_at_surface(const StackT<T> &up, const T *Rs, const T *sky, const T *gnd,
                                    T *G, T *X, T *vs, T *ur, T *dr, InvLds &sinv, int &np) {
  const int t = threadIdx.x;
  if (!inv_i_minus_ab<GJ>(up.Rp, Rs, X, G, sinv, np)) return false;
  T s_ = 0.0;
  if (t < NN) {
    s_ = mv_row(up.Rp, gnd, t);          // UPR (DNT INBOT)
    s_ = mv_row(up.Tp, sky, t) + s_;     // + UPT INTOP
    s_ = 0.0 + s_;                       // + UPR DNS (= 0)
    s_ = up.Sp[t] + s_;
  }
  __syncthreads();
  if (t < NN) vs[t] = s_;
  __syncthreads();
  if (t < NN) dr[t] = mv_row(G, vs, t);
  __syncthreads();
  if (!inv_i_minus_ab<GJ>(Rs, up.Rp, X, G, sinv, np)) return false;
  if (t < NN) vs[t] = mv_row(up.Tp, sky, t);
  __syncthreads();
  if (t < NN) {
    s_ = mv_row(Rs, vs, t);
    s_ = gnd[t] + s_;
    s_ = mv_row(Rs, up.Sp, t) + s_;
    s_ = 0.0 + s_;
  }
  __syncthreads();
  if (t < NN) vs[t] = s_;
  __syncthreads();
  if (t < NN) ur[t] = mv_row(G, vs, t);
  __syncthreads();
  return true;
}

// INTERNAL_RADIANCE4 at the top: up = nothing (R = 0, T = I, S = 0, both
// gammas the identity), down = stack including the surface
template <typename T>
__device__ void radiance_at_top(const StackT<T> &dn, const T *sky, const T *gnd, T *ur, T *dr) {
  const int t = threadIdx.x;
  if (t < NN) {
    dr[t] = sky[t];
    T s_ = mv_row(dn.Rm, sky, t);
    s_ = mv_row(dn.Tm, gnd, t) + s_;
    s_ = 0.0 + s_;
    ur[t] = dn.Sm[t] + s_;
  }
  __syncthreads();
}

// CONVERT_OUTPUT + collect_output (rt_utilities.f90:687-737,
// collect_output.f90:31-88) for one output level
template <typename T>
__device__ void write_tb(const Dims &d, const Settings &s, int col, int f, int io, double wl,
                         const T *ur, const T *dr, double *tb) {
  const int t = threadIdx.x;
  if (t < NMU) {
    const int j = t;
    double o_up[2], o_dn[2];
    const double iu = ur[NST * j], qu = ur[NST * j + 1];
    const double idn = dr[NST * j], qdn = dr[NST * j + 1];
    if (s.outpol == 0) {
      o_up[0] = rad2tb(2.0 * (0.5 * (iu + qu)), wl);
      o_up[1] = rad2tb(2.0 * (0.5 * (iu - qu)), wl);
      o_dn[0] = rad2tb(2.0 * (0.5 * (idn + qdn)), wl);
      o_dn[1] = rad2tb(2.0 * (0.5 * (idn - qdn)), wl);
    } else {
      o_up[0] = rad2tb(iu, wl);
      o_up[1] = rad2tb(qu, wl);
      o_dn[0] = rad2tb(idn, wl);
      o_dn[1] = rad2tb(qdn, wl);
    }
    const int su = NMU - 1 - j, sd = NMU + j;
    const size_t base = (size_t)col * d.nout + io;
    for (int k = 0; k < 2; k++) {
      tb[((base * PAM_NANGLES + su) * d.nf + f) * 2 + k] = o_up[k];
      tb[((base * PAM_NANGLES + sd) * d.nf + f) * 2 + k] = o_dn[k];
    }
  }
}

// The surface's reflection matrix R- (get_sfc_matrices, sfc_matrices.f90:88-194):
// specular/Fresnel 2x2 blocks on the diagonal, or the rank-one Lambertian
// 2 r mu_j w_j on the I component
template <typename T>
__device__ void surface_matrix(char refl, const double *Rf, const double *smu, const double *sqw,
                               T *Rs) {
  for (int e = threadIdx.x; e < NN2; e += RT_WG) {
    const int r = e / NN, c = e % NN;
    const int i2 = r % NST, j2 = r / NST, i1 = c % NST, j1 = c / NST;
    double v = 0.0;
    if (refl == 'L') {
      if (i2 == 0 && i1 == 0) v = 2.0 * Rf[0] * smu[j1] * sqw[j1];
    } else if (j1 == j2) {
      const double rv = Rf[j2], rh = Rf[NMU + j2];
      const double r1 = (rv + rh) / 2.0, r2 = (rv - rh) / 2.0;
      v = (i1 == i2) ? r1 : r2;
    }
    Rs[r * LD + c] = v;
  }
  __syncthreads();
}

// The radiances entering the column: sky (blackbody at SKY_TEMP, I only)
// at the top and the surface's emission at the bottom (thermal_radiance,
// surface.f90:169-195; get_sfc_matrices for the specular/Lambertian emission)
template <typename T>
__device__ void boundary_radiances(char refl, const double *E, const double *Rf, double psky,
                                   double pg, T *sky, T *gnd) {
  const int t = threadIdx.x;
  if (t < NN) {
    const int i = t % NST, j = t / NST;
    sky[t] = (i == 0) ? (1.0 - 0.0) * psky : 0.0;
    if (refl == 'L') {
      gnd[t] = (i == 0) ? E[0] * pg : 0.0;
    } else {
      const double rv = Rf[j], rh = Rf[NMU + j];
      const double r1 = (rv + rh) / 2.0, r2 = (rv - rh) / 2.0;
      gnd[t] = (i == 0) ? (1.0 - r1) * pg : -r2 * pg;
    }
  }
  __syncthreads();
}

// The adding kernels' matrices.  At N = 32 all of them live in LDS.  At
// N = 64 (33 KB each) only the stack R+, R-, T+, T- stays in LDS; the
// working matrices live in a global scratch of RT_GMATS matrices per
// workgroup slot (w.rt_scratch; the kernels then run at most one workgroup
// per slot, so blockIdx.x is the slot).  Generic pointers serve both.
constexpr int ADD_LDS = (NN == 32) ? 8 : 4;   // of the 8 matrices of k_adding
constexpr int LEV_LDS = (NN == 32) ? 10 : 4;  // of the 10 of k_adding_levels
static_assert(8 - ADD_LDS <= RT_GMATS && 10 - LEV_LDS <= RT_GMATS, "scratch per slot");

template <typename T>
__device__ __forceinline__ T *add_mat(T (*lds)[MSZ], int nlds, const Work &w, int k) {
  if (k < nlds) return lds[k];
  return reinterpret_cast<T *>(w.rt_scratch + ((size_t)blockIdx.x * RT_GMATS + (k - nlds)) * MSZ);
}

// In FP32 (the mixed path) all eight matrices live in LDS at either N
// (34 KB at N = 32, 133 KB at N = 64).
template <typename T> constexpr int add_lds() { return sizeof(T) == 4 ? 8 : ADD_LDS; }

template <typename T>
struct AddLdsT {
  T M[add_lds<T>()][MSZ];
  T v[10][NN];
  InvLds inv;
  double mu[NMU], qw[NMU];
  int bad[1];
  int slot;  // unit index of a persistent workgroup
};
typedef AddLdsT<double> AddLds;

// One (column, channel) unit of the chunk.  Returns false only in the
// series variant (GJ = false) when an inverse did not converge; the unit is
// then recomputed from scratch by the Gauss-Jordan pass (in FP64).
// T = float: the mixed path's adding, with the stack, the layer operators,
// the products, the inverse series and the radiances in FP32.
template <bool GJ, typename T = double>
__device__ bool adding_unit(const Dims &d, const Settings &s, const double *freqs,
                            const char *sfc_refl, const Work &w, double *tb, int cf,
                            AddLdsT<T> &LL) {
  constexpr int NL = add_lds<T>();
  T(*sv)[NN] = LL.v;
  InvLds &sinv = LL.inv;
  double *smu = LL.mu, *sqw = LL.qw;
  int *sbad = LL.bad;
  (void)s;

  const int t = threadIdx.x;
  const int cl = cf / d.nf, f = cf % d.nf;
  const int col = d.col0 + cl;
  if (t < NMU) {
    smu[t] = w.mu[t];
    sqw[t] = w.qw[t];
  }
  if (t == 0) sbad[0] = 0;
  int32_t *cst = w.cf_status + (size_t)col * d.nf + f;
  const int L = w.nl[col];
  const int nout = d.nout;
  const double wl = C_LIGHT / (freqs[f] * 1.e3);
  const char refl = sfc_refl ? sfc_refl[col] : 'S';
  __syncthreads();
  auto fail = [&](int32_t code) {
    if (t == 0 && code) atomicMax(cst, code);
    for (int e = t; e < nout * PAM_NANGLES * 2; e += RT_WG) {
      const int io = e / (PAM_NANGLES * 2), a = (e / 2) % PAM_NANGLES, k = e % 2;
      tb[((((size_t)col * nout + io) * PAM_NANGLES + a) * d.nf + f) * 2 + k] =
          __builtin_nan("");
    }
  };
  if ((w.col_status[col] & 0xff) >= PAM_FATAL) {
    fail(w.col_status[col]);
    return true;
  }
  if ((*cst & 0xff) >= PAM_FATAL) {
    fail(0);
    return true;
  }
  if (refl != 'S' && refl != 'F' && refl != 'L') {
    fail(E_REFL | PAM_WARNING);
    return true;
  }
  int lev_toa = -1, lev_sfc = -1;
  bool general = false;
  for (int io = 0; io < nout; io++) {
    const int o = max(1, min(w.outlev[(size_t)col * nout + io], L + 1));
    if (o == 1 && lev_toa < 0)
      lev_toa = io;
    else if (o == L + 1 && lev_sfc < 0)
      lev_sfc = io;
    else
      general = true;
  }
  if (general) {
    // a level inside the column (or a level asked for twice) needs the
    // stacks above and below it: the unit goes to k_adding_levels
    if (t == 0) {
      const unsigned q = atomicAdd(&w.redo_n[2], 1u);
      w.redo_lev[q] = cf;
    }
    return true;
  }
  StackT<T> stk = {add_mat(LL.M, NL, w, 0), add_mat(LL.M, NL, w, 1), add_mat(LL.M, NL, w, 2),
                   add_mat(LL.M, NL, w, 3), sv[0], sv[1]};
  T *R2 = add_mat(LL.M, NL, w, 4), *T2 = add_mat(LL.M, NL, w, 5);
  T *G = add_mat(LL.M, NL, w, 6), *X = add_mat(LL.M, NL, w, 7);
  T *S2p = sv[2], *S2m = sv[3], *dg = sv[4], *vs = sv[5];
  T *sky = sv[6], *gnd = sv[7], *ur = sv[8], *dr = sv[9];
  const double psky = planck(SKY_TEMP, wl);
  const double pg = planck(w.groundtemp[col], wl);
  const double *E = w.sfc_emis + ((size_t)col * d.nf + f) * 2 * NMU;
  const double *Rf = w.sfc_refl + ((size_t)col * d.nf + f) * 2 * NMU;
  boundary_radiances(refl, E, Rf, psky, pg, sky, gnd);
  // fold the layers top down (radtran4.f90:386-417 with prefix reuse: the
  // stack above the surface level is the stack the top level continues)
  unsigned long long nscat = 0, nclear = 0;
  int nprod = 0;
  // The clear layers above the highest scattering layer.  While the stack
  // holds only clear layers it is R = 0, T = diag(td), and the lanes of
  // wave 0 fold those layers in registers with the operations
  // combine_clear performs on the non-zero entries.  Lane k of each window
  // of 64 layers loads the bottom level of layer L-1-k and its Planck
  // radiance; the top level is the previous layer's bottom.
  int npre = L;
  for (int b = 0; b < L; b += 64) {
    const int k = b + (t & 63);
    const bool scat = k >= L || w.layer_item[(size_t)col * d.ML + (L - 1 - k)] >= 0;
    const unsigned long long m = __ballot(scat);
    if (m) {
      npre = min(L, b + __ffsll((long long)m) - 1);
      break;
    }
  }
  if (npre > 0) {
    if (t < 64) {
      const size_t ol = (size_t)col * (d.ML + 1);
      const double *kgc = w.kgas + ((size_t)col * d.nf + f) * d.ML;
      const int ii = t % NST;
      const double muj = smu[(t % NN) / NST];
      double hprev = w.hgt_lev[ol + L], pprev = planck(w.temp_lev[ol + L], wl);
      double td = 1.0, spv = 0.0, smv = 0.0;
      for (int b = 0; b < npre; b += 64) {
        const int k = b + t;
        double hk = 0.0, pk = 0.0, gk = 0.0;
        if (k < npre) {
          const int z = L - 1 - k;
          hk = w.hgt_lev[ol + z];
          pk = planck(w.temp_lev[ol + z], wl);
          gk = kgc[z];
        }
        const int cnt = min(64, npre - b);
        for (int q = 0; q < cnt; q++) {
          // NONSCATTER_LAYER4 (radintg4.f90:79-117), top hprev/pprev
          const double h0 = __shfl(hk, q, 64), p1 = __shfl(pk, q, 64), kg = __shfl(gk, q, 64);
          const double p0 = pprev;
          const double dtau = fabs(hprev - h0) * fmax(kg, 0.0);
          const double dgl = exp(-dtau / muj);
          double a = 0.0, bs = 0.0;
          if (ii == 0 && dtau > 0.0) {
            const double path = dtau / muj;
            const double slope = (p1 - p0) / path;
            a = p1 - slope - (p1 - slope * (1.0 + path)) * dgl;
            bs = p0 + slope - (p0 + slope * (1.0 + path)) * dgl;
          }
          if (b + q == 0) {
            td = dgl;
            spv = a;
            smv = bs;
          } else {
            smv = smv + td * bs;
            spv = a + dgl * spv;
            td = dgl * td;
          }
          hprev = h0;
          pprev = p1;
        }
      }
      if (t < NN) {
        stk.Sp[t] = spv;
        stk.Sm[t] = smv;
        dg[t] = td;
      }
    }
    __syncthreads();
    for (int e = t; e < NN2; e += RT_WG) {
      const int i = e / NN, j = e % NN;
      const double tt = (i == j) ? dg[i] : 0.0;
      stk.Rp[i * LD + j] = 0.0;
      stk.Rm[i * LD + j] = 0.0;
      stk.Tp[i * LD + j] = tt;
      stk.Tm[i * LD + j] = tt;
    }
    __syncthreads();
    nclear += (unsigned long long)(npre - 1);
  }
  LayerFetch F;
  if (npre < L) fetch_layer(d, w, col, f, L - 1 - npre, F);
  for (int l = npre + 1; l <= L; l++) {
    const int z = L - l;
    const bool sc = stage_layer(F, wl, smu, R2, T2, dg, S2p, S2m, sbad);
    if (l < L) fetch_layer(d, w, col, f, z - 1, F);
    if (l == 1) {
      for (int e = t; e < NN2; e += RT_WG) {
        const int i = e / NN, j = e % NN;
        const double r = sc ? R2[i * LD + j] : 0.0;
        const double tt = sc ? T2[i * LD + j] : (i == j ? dg[i] : 0.0);
        stk.Rp[i * LD + j] = r;
        stk.Rm[i * LD + j] = r;
        stk.Tp[i * LD + j] = tt;
        stk.Tm[i * LD + j] = tt;
      }
      if (t < NN) {
        stk.Sp[t] = S2p[t];
        stk.Sm[t] = S2m[t];
      }
      __syncthreads();
    } else if (sc) {
      if (!combine_scat<GJ>(stk, R2, T2, S2p, S2m, G, X, vs, sinv, nprod)) {
        if (!GJ) return false;
        fail(E_SINGULAR | PAM_FATAL);
        return true;
      }
      nscat++;
    } else {
      combine_clear(stk, dg, S2p, S2m);
      nclear++;
    }
  }
  if (sbad[0]) {
    fail(sbad[0]);
    return true;
  }
  surface_matrix(refl, Rf, smu, sqw, R2);
  if (s.fault == FAULT_SINGULAR && col == s.fault_col && f == s.fault_ch) {
    fail(E_SINGULAR | PAM_FATAL);  // test mode: the surface inverse failed
    return true;
  }
  if (lev_sfc >= 0) {
    if (!radiance_at_surface<GJ>(stk, R2, sky, gnd, G, X, vs, ur, dr, sinv, nprod)) {
      if (!GJ) return false;
      fail(E_SINGULAR | PAM_FATAL);
      return true;
    }
    write_tb(d, s, col, f, lev_sfc, wl, ur, dr, tb);
    __syncthreads();
  }
  if (lev_toa >= 0) {
    if (!combine_surface<GJ>(stk, R2, G, X, vs, sinv, nprod)) {
      if (!GJ) return false;
      fail(E_SINGULAR | PAM_FATAL);
      return true;
    }
    radiance_at_top(stk, sky, gnd, ur, dr);
    write_tb(d, s, col, f, lev_toa, wl, ur, dr, tb);
  }
  if (t == 0) {
    atomicAdd(&w.counters[4], nscat);
    atomicAdd(&w.counters[5], nclear);
    atomicAdd(&w.counters[10], (unsigned long long)nprod);
    if (GJ) atomicAdd(&w.counters[11], 1ull);
  }
  return true;
}

// T = float: the mixed path (at N = 32, 36 KB of LDS: four workgroups per
// compute unit, and the register budget that follows)
template <bool PERS, typename T>
__global__ void __launch_bounds__(RT_WG) __attribute__((amdgpu_waves_per_eu((NN == 32 && sizeof(T) == 4) ? 4 : 1))) k_adding(Dims d, Settings s, const double *freqs,
                                                  const char *sfc_refl, Work w, double *tb) {
  __shared__ AddLdsT<T> LL;
  const int n = d.ncol_chunk * d.nf;
  int cf = PERS ? next_unit(&w.queue[2], &LL.slot) : (int)blockIdx.x;
  while (cf < n) {
    // inverse_gj == 2 (a test mode) queues every unit for the Gauss-Jordan pass
    const bool done =
        s.inverse_gj != 2 && adding_unit<false, T>(d, s, freqs, sfc_refl, w, tb, cf, LL);
    if (!done && threadIdx.x == 0) {
      const unsigned q = atomicAdd(&w.redo_n[1], 1u);
      w.redo_add[q] = cf;
    }
    if (!PERS) break;
    cf = next_unit(&w.queue[2], &LL.slot);
  }
}

// Gauss-Jordan pass over the queued units (all = 0) or every unit (all = 1)
__global__ void __launch_bounds__(RT_WG) k_adding_gj(Dims d, Settings s, const double *freqs,
                                                     const char *sfc_refl, Work w, double *tb,
                                                     int all) {
  __shared__ AddLds LL;
  const int n = all ? d.ncol_chunk * d.nf : (int)w.redo_n[1];
  if (threadIdx.x == 0) LL.inv.gj_only = all;
  for (int q = blockIdx.x; q < n; q += gridDim.x) {
    __syncthreads();
    adding_unit<true>(d, s, freqs, sfc_refl, w, tb, all ? q : w.redo_add[q], LL);
  }
}

// INTERNAL_RADIANCE4 (radintg4.f90:119-172) between the stack above a level
// (its R+, T+, S+: URp, UTp, USp) and the stack below it (dn: R-, T-, S-),
// in the reference's order of the vector sums
__device__ bool radiance_internal(const double *URp, const double *UTp, const double *USp,
                                  const Stack &dn, const double *sky, const double *gnd,
                                  double *G, double *X, double *vs, double *ur, double *dr,
                                  InvLds &sinv, int &np) {
  const int t = threadIdx.x;
  if (!inv_i_minus_ab<true>(URp, dn.Rm, X, G, sinv, np)) return false;  // gamma+
  if (t < NN) vs[t] = mv_row(dn.Tm, gnd, t);                            // DNT INBOT
  __syncthreads();
  double s_ = 0.0;
  if (t < NN) {
    s_ = mv_row(URp, vs, t);
    s_ = mv_row(UTp, sky, t) + s_;
    s_ = mv_row(URp, dn.Sm, t) + s_;
    s_ = USp[t] + s_;
  }
  __syncthreads();
  if (t < NN) vs[t] = s_;
  __syncthreads();
  if (t < NN) dr[t] = mv_row(G, vs, t);
  __syncthreads();
  if (!inv_i_minus_ab<true>(dn.Rm, URp, X, G, sinv, np)) return false;  // gamma-
  if (t < NN) vs[t] = mv_row(UTp, sky, t);
  __syncthreads();
  if (t < NN) {
    s_ = mv_row(dn.Rm, vs, t);
    s_ = mv_row(dn.Tm, gnd, t) + s_;
    s_ = mv_row(dn.Rm, USp, t) + s_;
    s_ = dn.Sm[t] + s_;
  }
  __syncthreads();
  if (t < NN) vs[t] = s_;
  __syncthreads();
  if (t < NN) ur[t] = mv_row(G, vs, t);
  __syncthreads();
  return true;
}

// Folds the layers lo..hi (rt4's numbering, 1 = top) into stk, top down, the
// first one copied (radtran4.f90:386-417).  FP64 products; inverses by the
// series with the Gauss-Jordan fallback.  False: a singular matrix.
__device__ bool fold_layers(const Dims &d, const Work &w, int col, int f, int L, int lo, int hi,
                            double wl, const double *smu, Stack &stk, double *R2, double *T2,
                            double *dg, double *S2p, double *S2m, double *G, double *X, double *vs,
                            InvLds &sinv, int *sbad, int &np, unsigned long long &nscat,
                            unsigned long long &nclear) {
  const int t = threadIdx.x;
  LayerFetch F;
  fetch_layer(d, w, col, f, L - lo, F);
  for (int l = lo; l <= hi; l++) {
    const bool sc = stage_layer(F, wl, smu, R2, T2, dg, S2p, S2m, sbad);
    if (l < hi) fetch_layer(d, w, col, f, L - l - 1, F);
    if (l == lo) {
      for (int e = t; e < NN2; e += RT_WG) {
        const int i = e / NN, j = e % NN;
        const double r = sc ? R2[i * LD + j] : 0.0;
        const double tt = sc ? T2[i * LD + j] : (i == j ? dg[i] : 0.0);
        stk.Rp[i * LD + j] = r;
        stk.Rm[i * LD + j] = r;
        stk.Tp[i * LD + j] = tt;
        stk.Tm[i * LD + j] = tt;
      }
      if (t < NN) {
        stk.Sp[t] = S2p[t];
        stk.Sm[t] = S2m[t];
      }
      __syncthreads();
    } else if (sc) {
      if (!combine_scat<true>(stk, R2, T2, S2p, S2m, G, X, vs, sinv, np)) return false;
      nscat++;
    } else {
      combine_clear(stk, dg, S2p, S2m);
      nclear++;
    }
  }
  return true;
}

struct LevLds {
  double M[LEV_LDS][MSZ];
  double v[11][NN];
  InvLds inv;
  double mu[NMU], qw[NMU];
  int bad[1];
};

// One (column, channel) unit with an output level inside the column (or a
// level asked for twice).  As radtran4.f90:375-422 does for every level, each
// output level gets the stack of the layers above it and the stack of the
// layers below it with the surface, joined by INTERNAL_RADIANCE4.  FP64 in
// both precisions.  Output levels are 1..L+1 (run_rt.f90:166-184).
__device__ void levels_unit(const Dims &d, const Settings &s, const double *freqs,
                            const char *sfc_refl, const Work &w, double *tb, int cf, LevLds &LL) {
  double(*sv)[NN] = LL.v;
  const int t = threadIdx.x;
  const int cl = cf / d.nf, f = cf % d.nf;
  const int col = d.col0 + cl;
  if (t < NMU) {
    LL.mu[t] = w.mu[t];
    LL.qw[t] = w.qw[t];
  }
  if (t == 0) {
    LL.bad[0] = 0;
    LL.inv.gj_only = s.inverse_gj == 1;
  }
  int32_t *cst = w.cf_status + (size_t)col * d.nf + f;
  const int L = w.nl[col];
  const int nout = d.nout;
  const double wl = C_LIGHT / (freqs[f] * 1.e3);
  const char refl = sfc_refl ? sfc_refl[col] : 'S';
  __syncthreads();
  auto fail = [&](int32_t code) {
    if (t == 0 && code) atomicMax(cst, code);
    for (int e = t; e < nout * PAM_NANGLES * 2; e += RT_WG) {
      const int io = e / (PAM_NANGLES * 2), a = (e / 2) % PAM_NANGLES, k = e % 2;
      tb[((((size_t)col * nout + io) * PAM_NANGLES + a) * d.nf + f) * 2 + k] =
          __builtin_nan("");
    }
  };
  Stack stk = {add_mat(LL.M, LEV_LDS, w, 0), add_mat(LL.M, LEV_LDS, w, 1),
               add_mat(LL.M, LEV_LDS, w, 2), add_mat(LL.M, LEV_LDS, w, 3), sv[0], sv[1]};
  double *URp = add_mat(LL.M, LEV_LDS, w, 4), *UTp = add_mat(LL.M, LEV_LDS, w, 5);
  double *R2 = add_mat(LL.M, LEV_LDS, w, 6), *T2 = add_mat(LL.M, LEV_LDS, w, 7);
  double *G = add_mat(LL.M, LEV_LDS, w, 8), *X = add_mat(LL.M, LEV_LDS, w, 9);
  double *USp = sv[2], *S2p = sv[3], *S2m = sv[4], *dg = sv[5], *vs = sv[6];
  double *sky = sv[7], *gnd = sv[8], *ur = sv[9], *dr = sv[10];
  const double *E = w.sfc_emis + ((size_t)col * d.nf + f) * 2 * NMU;
  const double *Rf = w.sfc_refl + ((size_t)col * d.nf + f) * 2 * NMU;
  boundary_radiances(refl, E, Rf, planck(SKY_TEMP, wl), planck(w.groundtemp[col], wl), sky, gnd);
  unsigned long long nscat = 0, nclear = 0;
  int nprod = 0;
  for (int io = 0; io < nout; io++) {
    const int o = max(1, min(w.outlev[(size_t)col * nout + io], L + 1));
    // the layers above the level: none at the top (R = 0, T = I, S = 0)
    if (o > 1) {
      if (!fold_layers(d, w, col, f, L, 1, o - 1, wl, LL.mu, stk, R2, T2, dg, S2p, S2m, G, X, vs,
                       LL.inv, LL.bad, nprod, nscat, nclear)) {
        fail(E_SINGULAR | PAM_FATAL);
        return;
      }
    }
    for (int e = t; e < NN2; e += RT_WG) {
      const int i = e / NN, j = e % NN;
      URp[i * LD + j] = o > 1 ? stk.Rp[i * LD + j] : 0.0;
      UTp[i * LD + j] = o > 1 ? stk.Tp[i * LD + j] : (i == j ? 1.0 : 0.0);
    }
    if (t < NN) USp[t] = o > 1 ? stk.Sp[t] : 0.0;
    __syncthreads();
    // the layers below the level, then the surface (R+ = 0, R- = Rs, T = I,
    // S = 0); at the surface level the surface alone
    if (o <= L) {
      if (!fold_layers(d, w, col, f, L, o, L, wl, LL.mu, stk, R2, T2, dg, S2p, S2m, G, X, vs,
                       LL.inv, LL.bad, nprod, nscat, nclear)) {
        fail(E_SINGULAR | PAM_FATAL);
        return;
      }
      surface_matrix(refl, Rf, LL.mu, LL.qw, R2);
      if (!combine_surface<true>(stk, R2, G, X, vs, LL.inv, nprod)) {
        fail(E_SINGULAR | PAM_FATAL);
        return;
      }
    } else {
      surface_matrix(refl, Rf, LL.mu, LL.qw, stk.Rm);
      for (int e = t; e < NN2; e += RT_WG) {
        const int i = e / NN, j = e % NN;
        stk.Rp[i * LD + j] = 0.0;
        stk.Tp[i * LD + j] = i == j ? 1.0 : 0.0;
        stk.Tm[i * LD + j] = i == j ? 1.0 : 0.0;
      }
      if (t < NN) {
        stk.Sp[t] = 0.0;
        stk.Sm[t] = 0.0;
      }
      __syncthreads();
    }
    if (LL.bad[0]) {
      fail(LL.bad[0]);
      return;
    }
    if (!radiance_internal(URp, UTp, USp, stk, sky, gnd, G, X, vs, ur, dr, LL.inv, nprod)) {
      fail(E_SINGULAR | PAM_FATAL);
      return;
    }
    write_tb(d, s, col, f, io, wl, ur, dr, tb);
    __syncthreads();
  }
  if (t == 0) {
    atomicAdd(&w.counters[4], nscat);
    atomicAdd(&w.counters[5], nclear);
    atomicAdd(&w.counters[10], (unsigned long long)nprod);
  }
}

// The units queued by k_adding for output levels inside the column
__global__ void __launch_bounds__(RT_WG) k_adding_levels(Dims d, Settings s, const double *freqs,
                                                         const char *sfc_refl, Work w, double *tb) {
  __shared__ LevLds LL;
  const int n = (int)w.redo_n[2];
  for (int q = blockIdx.x; q < n; q += gridDim.x) {
    __syncthreads();
    levels_unit(d, s, freqs, sfc_refl, w, tb, w.redo_lev[q], LL);
  }
}


void launch_doubling(const Dims &d, const Settings &s, const double *freqs, Work &w,
                     hipStream_t st) {
  size_t n = (size_t)d.nitems * d.nf;
  if (n == 0) return;
  if (s.inverse_gj == 1) {
    hipLaunchKernelGGL(k_doubling_gj, dim3((unsigned)std::min<size_t>(n, 16384)), dim3(RT_WG), 0,
                       st, d, s, freqs, w, 1);
    return;
  }
  const dim3 grid((unsigned)persistent_grid(n, s.pers_dbl, s.ncu));
  const bool pers = grid.x < n;
  if (s.precision == PAM_PREC_MIXED && s.inverse_gj == 0) {
    if (pers)
      hipLaunchKernelGGL(k_doubling_mixed<true>, grid, dim3(RT_WG), 0, st, d, s, freqs, w);
    else
      hipLaunchKernelGGL(k_doubling_mixed<false>, grid, dim3(RT_WG), 0, st, d, s, freqs, w);
  } else {
    if (pers)
      hipLaunchKernelGGL(k_doubling<true>, grid, dim3(RT_WG), 0, st, d, s, freqs, w);
    else
      hipLaunchKernelGGL(k_doubling<false>, grid, dim3(RT_WG), 0, st, d, s, freqs, w);
  }
  hipLaunchKernelGGL(k_doubling_gj, dim3((unsigned)std::min<size_t>(n, 2048)), dim3(RT_WG), 0, st,
                     d, s, freqs, w, 0);
}

void launch_adding(const Dims &d, const Settings &s, const double *freqs, const char *sfc_refl,
                   Work &w, double *tb, hipStream_t st) {
  size_t n = (size_t)d.ncol_chunk * d.nf;
  if (n == 0) return;
  // at N = 64 the kernels use the global scratch of their slot: at most one
  // workgroup per CU, so blockIdx.x names the slot
  const size_t slots = RT_GMATS > 0 ? (size_t)std::max(1, s.ncu) : (size_t)1 << 30;
  if (s.inverse_gj == 1) {
    hipLaunchKernelGGL(k_adding_gj, dim3((unsigned)std::min<size_t>(n, std::min<size_t>(16384, slots))),
                       dim3(RT_WG), 0, st, d, s, freqs, sfc_refl, w, tb, 1);
  } else {
    dim3 grid((unsigned)persistent_grid(n, s.pers_add, s.ncu));
    if (RT_GMATS > 0) grid.x = (unsigned)std::min(n, slots);
    const bool pers = grid.x < n, mx = s.precision == PAM_PREC_MIXED;
    if (pers && mx)
      hipLaunchKernelGGL((k_adding<true, float>), grid, dim3(RT_WG), 0, st, d, s, freqs, sfc_refl, w,
                         tb);
    else if (pers)
      hipLaunchKernelGGL((k_adding<true, double>), grid, dim3(RT_WG), 0, st, d, s, freqs, sfc_refl, w,
                         tb);
    else if (mx)
      hipLaunchKernelGGL((k_adding<false, float>), grid, dim3(RT_WG), 0, st, d, s, freqs, sfc_refl, w,
                         tb);
    else
      hipLaunchKernelGGL((k_adding<false, double>), grid, dim3(RT_WG), 0, st, d, s, freqs, sfc_refl, w,
                         tb);
    hipLaunchKernelGGL(k_adding_gj, dim3((unsigned)std::min<size_t>(n, std::min<size_t>(2048, slots))),
                       dim3(RT_WG), 0, st, d, s, freqs, sfc_refl, w, tb, 0);
  }
  // units with output levels inside the column (none with the default levels)
  hipLaunchKernelGGL(k_adding_levels, dim3((unsigned)std::min<size_t>(n, std::min<size_t>(1024, slots))),
                     dim3(RT_WG), 0, st, d, s, freqs, sfc_refl, w, tb);
}

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

}  // namespace pam
