// pam_phase.hip -- the m = 0 phase matrix of every scattering (layer,
// channel) unit as matrix-core products.
//
// scatcnv (scatcnv.f90:103-192, rt_utilities.f90:242-493) evaluates, for
// each pair of quadrature angles and each of the NUMPTS/2+1 azimuths, the
// Legendre sums a1 = sum_l c1_l P_l(x), b1, a3 of the layer's scaled series
// and accumulates them with the rotation weights into e11, e12, e21, e22.
// Every entry is linear in the coefficients:
//   e11 = sum_l c1_l Q1_l,  e12 = sum_l c2_l Q12_l,  e21 = sum_l c2_l Q21_l,
//   e22 = sum_l (c1_l Qa_l - c3_l Qb_l),
// where the Q_l(pair) sum P_l(cos Theta_k) times the azimuth weight (and the
// rotation terms) over the azimuths; they depend only on NUMPTS and the
// angle pair.  So the Q are tabulated once per device for NUMPTS = 4, 6, ...,
// 204, and the units with equal NUMPTS are gathered into tiles of 16 whose
// phase matrices are one FP64 MFMA product each:
//   [16 units x (lmax+1) coefficients] x [(lmax+1) x 512 pairs] per entry.
// The two quadrants are written into the unit's R (-,+) and T (+,+) slots,
// where the doubling kernel reads them before it overwrites the slots with
// the layer's operators.
#include "pam_internal.h"

namespace pam {

constexpr int PH_WG = 256;
constexpr int PH_NPAIR = 2 * NMU * NMU;  // (quadrant, mu2, mu1)
constexpr int PH_TILE = 16;              // units per matrix-core tile
constexpr int PH_L1MAX = MAXNLEG + 1;

__host__ __device__ __forceinline__ int ph_lmax(int m) { return 2 * m; }

int64_t phase_table_layout(int64_t *off) {
  int64_t o = 0;
  for (int m = 0; m < PH_NBUCKET; m++) {
    off[m] = o;
    o += (int64_t)5 * (ph_lmax(m) + 1) * PH_NPAIR;
  }
  off[PH_NBUCKET] = o;
  return o;
}

// Legendre recurrence P_l = a_l x P_{l-1} - b_l P_{l-2}, a_l = (2l-1)/l,
// b_l = (l-1)/l
struct PhLeg {
  double a[MAXNLEG + 1], b[MAXNLEG + 1];
};
constexpr PhLeg make_phleg() {
  PhLeg r{};
  for (int l = 1; l <= MAXNLEG; l++) {
    r.a[l] = (2.0 * l - 1.0) / l;
    r.b[l] = (l - 1.0) / l;
  }
  return r;
}
__constant__ PhLeg c_phleg = make_phleg();

// Table of NUMPTS = 2m+4 at tab + off[m]: Q[c][l][pair] for the components
// c = Q1, Q12, Q21, Qa, Qb, l = 0..2m, scaled by 1/NUMPTS.  One thread per
// pair; the azimuths are summed in the reference's order.
__global__ void __launch_bounds__(PH_WG) k_phase_tables(double *tab, const int64_t *off,
                                                        const double *mu) {
  const int m = blockIdx.x, p = blockIdx.y * PH_WG + threadIdx.x;
  const int lmax = ph_lmax(m), L1 = lmax + 1, numpts = 2 * m + 4;
  double *Q = tab + off[m];
  const int q = p / (NMU * NMU), j2 = (p / NMU) % NMU, j1 = p % NMU;
  const double mu1 = q ? -mu[j1] : mu[j1];
  const double mu2 = mu[j2];
  const double st1 = sqrt(1. - mu1 * mu1), st2 = sqrt(1. - mu2 * mu2);
  const double sq12 = sqrt((1. - mu1 * mu1) * (1. - mu2 * mu2));
  for (int c = 0; c < 5; c++)
    for (int l = 0; l < L1; l++) Q[((size_t)c * L1 + l) * PH_NPAIR + p] = 0.0;
  for (int k = 1; k <= numpts / 2 + 1; k++) {
    const double delphi = (2.0 * PI * (k - 1)) / numpts;
    const double cosphi = cos(delphi), sinphi = sin(delphi);
    const double x = mu1 * mu2 + sq12 * cosphi;
    const double sin_scat = sqrt(fmax(0.0, 1. - x * x));
    double sin1, sin2, cos1, cos2;
    if (sin_scat == 0.0) {
      sin1 = 0.0;
      sin2 = 0.0;
      cos1 = 1.0;
      cos2 = -1.0;
    } else {
      const double rs = 1.0 / sin_scat;
      sin1 = st2 * sinphi * rs;
      sin2 = st1 * sinphi * rs;
      cos1 = (st1 * mu2 - st2 * mu1 * cosphi) * rs;
      cos2 = (st2 * mu1 - st1 * mu2 * cosphi) * rs;
    }
    const double sin21 = 2.0 * sin1 * cos1, cos21 = 1.0 - 2.0 * sin1 * sin1;
    const double sin22 = 2.0 * sin2 * cos2, cos22 = 1.0 - 2.0 * sin2 * sin2;
    const double wk = (k == 1 || k == numpts / 2 + 1) ? 1.0 : 2.0;
    const double w0 = wk, w1 = wk * cos21, w2 = wk * cos22, w3 = wk * (cos21 * cos22),
                 w4 = wk * (sin21 * sin22);
    double pl = 1.0, pl1 = 1.0, pl2 = 0.0;
    for (int l = 0; l <= lmax; l++) {
      if (l > 0) pl = (c_phleg.a[l] * x) * pl1 - c_phleg.b[l] * pl2;
      Q[((size_t)0 * L1 + l) * PH_NPAIR + p] += w0 * pl;
      Q[((size_t)1 * L1 + l) * PH_NPAIR + p] += w1 * pl;
      Q[((size_t)2 * L1 + l) * PH_NPAIR + p] += w2 * pl;
      Q[((size_t)3 * L1 + l) * PH_NPAIR + p] += w3 * pl;
      Q[((size_t)4 * L1 + l) * PH_NPAIR + p] += w4 * pl;
      pl2 = pl1;
      pl1 = pl;
    }
  }
  for (int c = 0; c < 5; c++)
    for (int l = 0; l < L1; l++) Q[((size_t)c * L1 + l) * PH_NPAIR + p] /= numpts;
}

void build_phase_tables(double *tab, const int64_t *d_off, const double *d_mu, hipStream_t st) {
  hipLaunchKernelGGL(k_phase_tables, dim3(PH_NBUCKET, PH_NPAIR / PH_WG), dim3(PH_WG), 0, st, tab,
                     d_off, d_mu);
}

// ---- gathering the units by NUMPTS (a counting sort)

__device__ __forceinline__ bool ph_scatters(const Settings &s, const Work &w, int itf) {
  return s.lhyd_scat && !(w.itf_status[itf] & 0xff) && w.nleg[itf] > 0;
}

__device__ __forceinline__ int ph_bucket(int nleg) { return (2 * ((nleg + 1) / 2)) / 2; }

// ranks within the buckets: a histogram per workgroup in LDS, then one
// global atomic per bucket and workgroup
__global__ void __launch_bounds__(256) k_phase_count(int nif, Settings s, Work w) {
  __shared__ uint32_t hist[PH_NBUCKET], base[PH_NBUCKET];
  const int t = threadIdx.x;
  for (int b = t; b < PH_NBUCKET; b += blockDim.x) hist[b] = 0;
  __syncthreads();
  const int itf = blockIdx.x * blockDim.x + t;
  const bool on = itf < nif && ph_scatters(s, w, itf);
  const int m = on ? ph_bucket(w.nleg[itf]) : 0;
  const uint32_t local = on ? atomicAdd(&hist[m], 1u) : 0u;
  __syncthreads();
  for (int b = t; b < PH_NBUCKET; b += blockDim.x)
    base[b] = hist[b] ? atomicAdd(&w.ph_cnt[b], hist[b]) : 0u;
  __syncthreads();
  if (on) w.ph_rank[itf] = (int32_t)(base[m] + local);
}

// offsets of the buckets (units) and of their tiles; ph_cnt[PH_NBUCKET]
// receives the number of tiles
__global__ void k_phase_scan(Work w) {
  if (threadIdx.x != 0) return;
  uint32_t o = 0, to = 0;
  for (int m = 0; m < PH_NBUCKET; m++) {
    w.ph_off[m] = o;
    w.ph_toff[m] = to;
    o += w.ph_cnt[m];
    to += (w.ph_cnt[m] + PH_TILE - 1) / PH_TILE;
  }
  w.ph_off[PH_NBUCKET] = o;
  w.ph_toff[PH_NBUCKET] = to;
}

__global__ void k_phase_scatter(int nif, Settings s, Work w) {
  const int itf = blockIdx.x * blockDim.x + threadIdx.x;
  if (itf >= nif) return;
  if (!ph_scatters(s, w, itf)) return;
  const int m = ph_bucket(w.nleg[itf]);
  w.ph_perm[w.ph_off[m] + w.ph_rank[itf]] = itf;
}

typedef double d4p_t __attribute__((ext_vector_type(4)));
typedef float f4p_t __attribute__((ext_vector_type(4)));

// The matrix-core step of k_phase per element type.  FP64 (16x16x4 f64):
// lane l supplies A[l%16][l/16], B[l/16][l%16] and holds D[4*r+l/16][l%16],
// so its result r belongs to the tile's unit 4 r + l/16.  FP32 (16x16x4
// f32): lane l holds D[4*(l/16)+r][l%16], unit 4 (l/16) + r.
template <typename R> struct PhMfma;
template <> struct PhMfma<double> {
  typedef d4p_t acc_t;
  static __device__ __forceinline__ acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int unit(int lk, int r) { return 4 * r + lk; }
};
template <> struct PhMfma<float> {
  typedef f4p_t acc_t;
  static __device__ __forceinline__ acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int unit(int lk, int r) { return 4 * lk + r; }
};

// One tile: up to 16 units of one NUMPTS.  Each wave computes 16 of the 64
// output tile pairs (16 angle pairs, entries e11/e12 or e21/e22).  R = double:
// the FP64 tables and phase matrix; R = float (the mixed path): the FP32 copy
// of the tables, FP32 products and an FP32 phase matrix, half the bytes read
// and written.
template <typename R>
__global__ void __launch_bounds__(PH_WG) k_phase(Work w) {
  typedef PhMfma<R> MF;
  typedef typename MF::acc_t acc_t;
  __shared__ R sA[3][PH_TILE][PH_L1MAX + 3];  // scaled series of the tile's units
  __shared__ int s_unit[PH_TILE];
  __shared__ int s_m, s_first, s_count;
  const int t = threadIdx.x, wv = t >> 6, lane = t & 63, li = lane & 15, lk = lane >> 4;
  const int tile = blockIdx.x;
  if (t == 0) {
    s_m = -1;
    const uint32_t ntiles = w.ph_toff[PH_NBUCKET];
    if ((uint32_t)tile < ntiles) {
      int m = 0;
      while (w.ph_toff[m + 1] <= (uint32_t)tile) m++;
      const int k = tile - (int)w.ph_toff[m];
      s_m = m;
      s_first = (int)w.ph_off[m] + k * PH_TILE;
      s_count = min(PH_TILE, (int)w.ph_cnt[m] - k * PH_TILE);
    }
  }
  __syncthreads();
  const int m = s_m;
  if (m < 0) return;
  const int lmax = ph_lmax(m), L1 = lmax + 1, L1p = (L1 + 3) & ~3;
  if (t < PH_TILE) s_unit[t] = t < s_count ? w.ph_perm[s_first + t] : -1;
  __syncthreads();
  // coefficients: c_l * kext * salb / (4 pi), zero above the unit's nlegen
  for (int e = t; e < 3 * PH_TILE * L1p; e += PH_WG) {
    const int c = e / (PH_TILE * L1p), i = (e / L1p) % PH_TILE, l = e % L1p;
    const int itf = s_unit[i];
    double v = 0.0;
    if (itf >= 0) {
      const int nleg = w.nleg[itf];
      if (l <= nleg && l < L1) {
        const double cst = w.kext[itf] * w.salb[itf] / (4.0 * 3.1415926535897932384);
        v = cst * w.coef[(size_t)itf * NCOEF * MAXNLEG + c * MAXNLEG + l];
      }
    }
    sA[c][i][l] = (R)v;
  }
  __syncthreads();
  const R *Q = (sizeof(R) == 4 ? (const R *)w.ptab_f : (const R *)w.ptab) + w.ptab_off[m];
  // entries (e11, e12) and (e21, e22) of 16 pairs per iteration, so that a
  // lane stores two adjacent doubles of its row
  for (int T = wv; T < 2 * (PH_NPAIR / 16); T += 4) {
    const int hi = T / (PH_NPAIR / 16), pt = T % (PH_NPAIR / 16);
    const int pair = 16 * pt + li;
    acc_t acc0 = {(R)0, (R)0, (R)0, (R)0}, acc1 = acc0;
    // hi = 0: e11 = c1 Q1, e12 = c2 Q12; hi = 1: e21 = c2 Q21, e22 = c1 Qa - c3 Qb
    const int ca0 = hi ? 1 : 0, qa0 = hi ? 2 : 0, ca1 = hi ? 0 : 1, qa1 = hi ? 3 : 1;
    for (int sk = 0; sk < L1p / 4; sk++) {
      const int l = 4 * sk + lk;
      const bool in = l < L1;
      const R b0 = in ? Q[((size_t)qa0 * L1 + l) * PH_NPAIR + pair] : (R)0;
      const R b1 = in ? Q[((size_t)qa1 * L1 + l) * PH_NPAIR + pair] : (R)0;
      acc0 = MF::step(sA[ca0][li][l], b0, acc0);
      acc1 = MF::step(sA[ca1][li][l], b1, acc1);
    }
    if (hi) {  // e22: - sum_l c3_l Qb_l
      for (int sk = 0; sk < L1p / 4; sk++) {
        const int l = 4 * sk + lk;
        const R b = l < L1 ? -Q[((size_t)4 * L1 + l) * PH_NPAIR + pair] : (R)0;
        acc1 = MF::step(sA[2][li][l], b, acc1);
      }
    }
    const int q = pair / (NMU * NMU), j2 = (pair / NMU) % NMU, j1 = pair % NMU;
    const int r = NST * j2 + hi, c = NST * j1;
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      const int itf = s_unit[MF::unit(lk, rr)];
      if (itf >= 0) {
        double *slot = (q ? w.R : w.T) + (size_t)itf * NN2;
        if constexpr (sizeof(R) == 4)
          *reinterpret_cast<float2 *>(reinterpret_cast<float *>(slot) + r * NN + c) =
              make_float2(acc0[rr], acc1[rr]);
        else
          *reinterpret_cast<double2 *>(slot + r * NN + c) = make_double2(acc0[rr], acc1[rr]);
      }
    }
  }
}

void launch_phase(const Dims &d, const Settings &s, Work &w, hipStream_t st) {
  const int nif = d.nitems * d.nf;
  if (nif == 0 || !s.lhyd_scat) return;
  launch_fill_u32(w.ph_cnt, PH_NBUCKET + 1, 0, st);
  const int nb = (nif + 255) / 256;
  hipLaunchKernelGGL(k_phase_count, dim3(nb), dim3(256), 0, st, nif, s, w);
  hipLaunchKernelGGL(k_phase_scan, dim3(1), dim3(64), 0, st, w);
  hipLaunchKernelGGL(k_phase_scatter, dim3(nb), dim3(256), 0, st, nif, s, w);
  const int ntile_max = nif / PH_TILE + PH_NBUCKET;
  if (s.precision == PAM_PREC_MIXED && w.ptab_f)  // as phase_in_f32
    hipLaunchKernelGGL(k_phase<float>, dim3(ntile_max), dim3(PH_WG), 0, st, w);
  else
    hipLaunchKernelGGL(k_phase<double>, dim3(ntile_max), dim3(PH_WG), 0, st, w);
}

}  // namespace pam
