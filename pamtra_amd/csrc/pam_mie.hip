// pam_mie.hip -- Mie single scattering of a layer's hydrometeors and the
// Legendre expansion of the layer phase function: calc_mie_spheres
// (mie_spheres.f90:14-345) with miecalc/miecross/mieangle
// (mie_scat_utilities.f90:8-160), accumulated over hydrometeors as
// calc_scatProperties does (scatProperties.f90:920-946) and normalised as in
// finalize_rt3_scatProperties (:993-1002).
//
// One 256-thread workgroup per (scattering layer, channel), three phases per
// hydrometeor:
//   A. threads over size bins: the a_n, b_n recurrences (miecalc) and the
//      efficiencies (miecross); coefficients staged in LDS, bins in chunks;
//   B. threads over (quadrature angle, bin group): S1/S2 and the DSD-weighted
//      phase-function sums (mieangle), pi_n/tau_n read from the angle table;
//   C. threads over Legendre orders: the projection sum over the angles in
//      ascending order, as the reference sums it.
// The Gauss-Legendre nodes, pi_n(mu_k), tau_n(mu_k) and P_l(mu_k) depend
// only on the series length N (nquad and nlegen are functions of N), so they
// are tabulated once per context (build_angle_tables) with the recurrences
// the reference evaluates per bin.
#include <algorithm>

#include "pam_internal.h"

namespace pam {

constexpr int MIE_WG = 256;
static_assert(MAX_NQUAD <= MIE_WG && MAXNLEG <= MIE_WG, "one thread per angle / order");

// nlegen and nquad of a series of length N (mie_spheres.f90:139-158)
__host__ __device__ __forceinline__ int mie_nlegen(int n) {
  return (2 * n > MAXNLEG - 1) ? MAXNLEG - 1 : 2 * n;
}
__host__ __device__ __forceinline__ int mie_nquad(int n) { return (mie_nlegen(n) + 2 * n + 2) / 2; }

// TKC liquid water permittivity (eps_water.f90:653-757), T in degC, f GHz;
// rounded to single precision like the reference's default-kind cmplx()
__device__ cd eps_water_tkc(double T, double f) {
  const double a1 = 8.169396e+01, b1 = 4.410555e-03, c1 = 1.208992e-13, d1 = 6.768869e+02;
  const double a2 = 1.597733e+00, b2 = 1.060228e-02, c2 = 9.982113e-15, d2 = 5.720517e+02;
  const double tc = 1.351758e+02;
  double frq = f * 1e9;
  double eps_s = 87.9144 - 0.404399 * T + 9.58726e-4 * (T * T) - 1.32802e-6 * (T * T * T);
  double delta1 = a1 * exp(-b1 * T), tau1 = c1 * exp(d1 / (T + tc));
  double delta2 = a2 * exp(-b2 * T), tau2 = c2 * exp(d2 / (T + tc));
  double w = 2.0 * PI * frq;
  double x1 = w * tau1, x2 = w * tau2;
  double t1 = (tau1 * tau1 * delta1) / (1.0 + x1 * x1);
  double t2 = (tau2 * tau2 * delta2) / (1.0 + x2 * x2);
  double e1 = eps_s - (w * w) * (t1 + t2);
  t1 = (tau1 * delta1) / (1.0 + x1 * x1);
  t2 = (tau2 * delta2) / (1.0 + x2 * x2);
  double e2 = w * (t1 + t2);
  return cd{sp(e1), sp(e2)};
}

// ice permittivity (eps_ice.f90:1-55), T in K
__device__ cd eps_ice(double T, double f) {
  double er = (T >= 240.0) ? 3.1884 + 9.1e-4 * (T - 273.0) : 3.1884 + 9.1e-4 * (240.0 - 273.0);
  double mit = (300.0 / T) - 1.0;
  double alpha = (0.00504 + 0.0062 * mit) * exp(-22.1 * mit);
  double e335 = exp(335.0 / T);
  double beta_m = (0.0207 / T) * e335 / ((e335 - 1.0) * (e335 - 1.0));
  beta_m = beta_m + 1.16e-11 * (f * f);
  double delta_beta = exp(-9.963 + 0.0372 * (T - 273.16));
  return cd{er, alpha / f + (beta_m + delta_beta) * f};
}

__device__ __forceinline__ int mie_nstop(double x) {
  return (int)(x + 4.0 * pow(x, sp(0.3334)) + 2.0);
}

// Angle table of series length N at tab + off[N]:
//   mu[Q], wt[Q]          Gauss-Legendre nodes (ascending) and weights
//   pi[N][Q], tau[N][Q]   c_n pi_n(mu_k), c_n tau_n(mu_k), n = 1..N, with
//                         c_n = (2n+1)/(n(n+1)) in single precision
//                         (mie_scat_utilities.f90:144)
//   P[Q][nlegen+1]        P_l(mu_k), l = 0..nlegen
int64_t angle_table_layout(int64_t *off, int *nmax) {
  int64_t o = 0;
  int n = 1;
  off[0] = 0;
  for (; n <= MAX_NTERMS; n++) {
    const int q = mie_nquad(n), nl = mie_nlegen(n);
    if (q > MAX_NQUAD) break;
    off[n] = o;
    o += (int64_t)q * (2 + 2 * n + nl + 1);
  }
  *nmax = n - 1;
  return o;
}

__global__ void __launch_bounds__(MIE_WG) k_angle_tables(double *tab, const int64_t *off) {
  const int n = blockIdx.x + 1, t = threadIdx.x;
  const int q = mie_nquad(n), nl = mie_nlegen(n);
  double *mu = tab + off[n], *wt = mu + q, *pit = wt + q, *taut = pit + (size_t)n * q,
         *pt = taut + (size_t)n * q;
  // gausquad (scat_utilities.f90:1-44)
  for (int j = t + 1; j <= (q + 1) / 2; j += MIE_WG) {
    double x = cos(PI * (j - 0.25) / (q + 0.5));
    double pl, pl1, pl2, dpl, xp;
    int it = 0;
    do {
      pl1 = 1.0;
      pl = x;
      for (int l = 2; l <= q; l++) {
        pl2 = pl1;
        pl1 = pl;
        pl = ((double)(2 * l - 1) * x * pl1 - (double)(l - 1) * pl2) / (double)l;
      }
      dpl = (double)q * (x * pl - pl1) / (x * x - 1.0);
      xp = x;
      x = xp - pl / dpl;
      it++;
    } while (fabs(x - xp) > 3.0e-13 && it < 10);
    const double w = 2.0 / ((1.0 - x * x) * dpl * dpl);
    mu[j - 1] = -x;
    wt[j - 1] = w;
    mu[q - j] = x;
    wt[q - j] = w;
  }
  __syncthreads();
  if (t < q) {
    const double m = mu[t];
    // pi_n, tau_n (mie_scat_utilities.f90:128-160)
    double pin = 1.0, pim = 0.0;
    for (int k = 1; k <= n; k++) {
      const double taun = (double)k * m * pin - (double)(k + 1) * pim;
      const double cn = (double)((float)(2 * k + 1) / (float)(k * (k + 1)));
      pit[(size_t)(k - 1) * q + t] = cn * pin;
      taut[(size_t)(k - 1) * q + t] = cn * taun;
      const double tmp = pin;
      pin = ((double)(2 * k + 1) * m * pin - (double)(k + 1) * pim) / (double)k;
      pim = tmp;
    }
    // Legendre polynomials of the projection (mie_spheres.f90:272-314)
    double pl = 1.0, pl1 = 1.0, pl2 = 0.0;
    for (int l = 0; l <= nl; l++) {
      if (l > 0) pl = (double)(2 * l - 1) * m * pl1 / (double)l - (double)(l - 1) * pl2 / (double)l;
      pt[(size_t)t * (nl + 1) + l] = pl;
      pl2 = pl1;
      pl1 = pl;
    }
  }
}

void build_angle_tables(double *tab, const int64_t *d_off, int nmax, hipStream_t st) {
  hipLaunchKernelGGL(k_angle_tables, dim3((unsigned)nmax), dim3(MIE_WG), 0, st, tab, d_off);
}

__global__ void k_round_f32(const double *src, float *dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

void round_tables_f32(const double *tab, float *tab_f, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_round_f32, dim3(1024), dim3(256), 0, st, tab, tab_f, n);
}

// Phase B as a matrix product on the FP64 matrix cores.  For a chunk of Bc
// bins the coefficient matrix U (2 Bc rows: real and imaginary parts; Kp =
// 2 Np columns: c_n a_n for n < Np, then c_n b_n) times
//   P1 = [pi; tau] and P2 = [tau; pi]   (Kp x nquad, from the angle table)
// gives Re/Im of S1 and S2 at every Gauss angle: S1 = U P1, S2 = U P2.
// U rows are ordered per 16-row tile as [re b0..b3, im b0..b3, re b4..b7,
// im b4..b7], so that each lane's MFMA result holds the real and imaginary
// parts of two bins at one angle.
constexpr int U_CAP = 4096;       // doubles of U staged in LDS per chunk
constexpr int MIE_MAXBIN = 512;   // size bins per hydrometeor supported by the chunk plan
constexpr int MIE_MAXCHUNK = 160;

__device__ __forceinline__ int u_row_re(int j) { return 16 * (j >> 3) + 8 * ((j & 7) >> 2) + (j & 3); }

// 1/x from the hardware reciprocal and two Newton steps (to ~1 ulp)
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// a / b for the Mie recurrences, with one reciprocal of |b|^2 (the operands
// stay far from the overflow range that Smith's algorithm guards against)
__device__ __forceinline__ cd cdiv_m(cd a, cd b) {
  const double r = rcp_nr(b.re * b.re + b.im * b.im);
  return cd{(a.re * b.re + a.im * b.im) * r, (a.im * b.re - a.re * b.im) * r};
}

typedef double d4m_t __attribute__((ext_vector_type(4)));
typedef float f4m_t __attribute__((ext_vector_type(4)));

// LDS stage of the angle table (each of c_n pi_n and c_n tau_n, series
// length x angles with mu >= 0); longer tables are read from memory.
constexpr int MIE_TAB = 1488;

// Sum over the 64 lanes of a wave by a fixed butterfly (the same order on
// every call, so the result is deterministic).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Column of U and series index n (1-based) of position k of a parity group
// (group 0: a_n for odd n, then b_n for even n; group 1: a_n for even n,
// then b_n for odd n); `apart` tells an a-column.
__device__ __forceinline__ void mie_kmap(int g, int k, int np, int &colu, int &n, bool &apart) {
  const int h = np >> 1;
  apart = k < h;
  const int m = apart ? k : k - h;
  n = apart ? 2 * m + 1 + g : 2 * m + 2 - g;
  colu = apart ? n - 1 : np + n - 1;
}

struct MieLds {
  double u[U_CAP];             // U of the chunk; at the end the normalised phase sums
  double bin[3][MIE_WG];       // per bin of the chunk: conj(m) (re, im) and 1/x
  double wb[MIE_WG];           // per-bin DSD weight n(D) dD of the chunk, 0 if skipped
  double ps[3][MAX_NQUAD];     // weighted phase sums per angle
  double tab[2 * MIE_TAB];     // angle table stage
  double pw[4][6 * 16];        // per-wave phase sums of shared angle tiles
  double lc[NCOEF][MAXNLEG + 1];
  double red[4][3];            // per-wave efficiency sums of the chunk
  int redi[4][2];              // per-wave bin and series counts
  double hm[MAX_HYDRO][2];     // per hydrometeor: refractive index
  int hnb[MAX_HYDRO], hnt[MAX_HYDRO];  // bins, series length (0: none)
  short nmb[MIE_MAXBIN];       // series length of every bin (0: skipped)
  short cb[MIE_MAXCHUNK + 1], cnp[MIE_MAXCHUNK], cnm[MIE_MAXCHUNK];
  int nc;
  double sum[3];
  int nlh, nan_;
  int slot;                    // unit index of a persistent workgroup
};

// ---------------------------------------------------------------------
// Self-similar Rayleigh-Gans aggregates in Legendre-series form
// (calc_self_similar_rayleigh_gans_rt3, rayleigh_gans.f90:467-846): the
// same outputs as the Mie spheres (extinction, albedo, backscattering and
// the normalised phase-function sums at the Gauss nodes), so the layer
// accumulation, the Legendre projection and the RT are shared.
// ---------------------------------------------------------------------
constexpr int SSRG_QANG = 360;    // angles of the scattering integral, 0..pi
constexpr int SSRG_TG = U_CAP - 3 * MIE_MAXBIN;  // (2j)^-gamma table in the U stage
static_assert(SSRG_TG >= 1024, "power table");

// shape factor of the SSRGA phase function at electrical size x
// (calc_shape_factor, rayleigh_gans.f90:848-872); tg[j-1] = (2j)^-gamma
__device__ double ssrg_shape(double x, double kappa, double beta, double gamma, double zeta,
                             const double *tg) {
  // tg: (2j)^-gamma for j <= SSRG_TG, or null (per-bin gamma): computed here
  const double c = cos(x);
  const double a = (1.0 + kappa / 3.0) * (1.0 / (2.0 * x + PI) - 1.0 / (2.0 * x - PI)) -
                   kappa * (1.0 / (2.0 * x + 3.0 * PI) - 1.0 / (2.0 * x - 3.0 * PI));
  const double scale = (c * a) * (c * a);
  const int jmax = (int)floor(5.0 * x / PI + 1.0);
  double tp = 2.0 * (x + PI), tm = 2.0 * (x - PI);
  double summ = zeta * (tg ? tg[0] : pow(2.0, -gamma)) * (1.0 / (tp * tp) + 1.0 / (tm * tm));
  for (int j = 2; j <= jmax; j++) {
    const double g = (tg && j <= SSRG_TG) ? tg[j - 1] : pow(2.0 * j, -gamma);
    tp = 2.0 * (x + PI * j);
    tm = 2.0 * (x - PI * j);
    // 1/tp^2 + 1/tm^2 with one refined reciprocal instead of two divisions
    const double a2 = tp * tp, b2 = tm * tm;
    summ = summ + g * ((a2 + b2) * rcp_nr(a2 * b2));
  }
  const double sx = sin(x);
  summ = summ * beta * (sx * sx);
  return PI * PI * 0.25 * (scale + summ);
}

// Efficiency sums and phase-function sums of one SSRGA hydrometeor.  Waves
// take bins for the 359-angle scattering integral (the per-bin terms summed
// afterwards in bin order, as the reference does); threads take the Gauss
// nodes for the phase-function sums, bins in order.  Returns false after
// setting the status on an input the reference refuses.
// Per-bin inputs: the descriptor's values, or the full spectrum's lists
// (scatProperties.f90:258-287: negative or NaN canting and as_ratio read as 0).
struct SsrgBins {
  const double *ar, *cant, *kap, *bet, *gam, *zet;  // per bin, or null
  double ar0, cant0, kap0, bet0, gam0, zet0;
  __device__ static double clean(double v) { return (v < 0.0 || isnan(v)) ? 0.0 : v; }
  __device__ double as_ratio(int i) const { return clean(ar ? ar[i] : ar0); }
  __device__ double canting(int i) const { return clean(cant ? cant[i] : cant0); }
  __device__ double kappa(int i) const { return kap ? kap[i] : kap0; }
  __device__ double beta(int i) const { return bet ? bet[i] : bet0; }
  __device__ double gamma(int i) const { return gam ? gam[i] : gam0; }
  __device__ double zeta(int i) const { return zet ? zet[i] : zet0; }
};

__device__ bool ssrg_sums(const pam_hydro_desc &hd, const SsrgBins &pb, int nbin,
                          const double *dia, const double *dd, const double *nd,
                          const double *mass, cd refr, double wavelength, int nquad,
                          const double *tmu, double *s_u, double (*s_ps)[MAX_NQUAD], double *sumq,
                          int *s_flag, int32_t *st) {
  const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
  // assertions of rayleigh_gans.f90:590-616, over the bins
  if (t == 0) *s_flag = (hd.liq_ice != -1);
  __syncthreads();
  for (int i = t; i < nbin; i += MIE_WG) {
    const double c = pb.canting(i);
    if (!(c == 0.0 || c == 90.0) || !(pb.as_ratio(i) > 0.0)) *s_flag = 1;
  }
  __syncthreads();
  if (*s_flag) {
    if (t == 0) set_err(st, E_SSRG_INPUT);
    return false;
  }
  double *bq = s_u, *tg = pb.gam ? nullptr : s_u + 3 * MIE_MAXBIN;
  if (tg)
    for (int j = t; j < SSRG_TG; j += MIE_WG) tg[j] = pow(2.0 * (j + 1), -pb.gam0);
  // dielectric factor of solid ice, (n + i k)^2 (rayleigh_gans.f90:642-645)
  const cd eps = refr * refr;
  const cd K = cdiv(eps - 1.0, eps + 2.0);
  const double K2 = K.re * K.re + K.im * K.im;
  const double k = 2.0 * PI / wavelength;
  const double k2 = k * k;
  __syncthreads();
  // scattering integral every 1/359 of pi (rayleigh_gans.f90:713-725)
  constexpr int LAST_LANE = (SSRG_QANG - 2) % 64;
  for (int ii = wv; ii < nbin; ii += MIE_WG / 64) {
    if (!(nd[ii] > 0.0)) {
      if (lane == 0) bq[ii] = bq[MIE_MAXBIN + ii] = bq[2 * MIE_MAXBIN + ii] = 0.0;
      continue;
    }
    const double dw = (pb.canting(ii) == 0.0) ? dia[ii] * pb.as_ratio(ii) : dia[ii];
    const double vol = mass[ii] / 917.0;
    const double pref = 9.0 * (k2 * k2) * K2 * (vol * vol) / (4.0 * PI);
    const double kappa = pb.kappa(ii), beta = pb.beta(ii), gamma = pb.gamma(ii), zeta = pb.zeta(ii);
    double q = 0.0, last = 0.0;
    for (int ia = 2 + lane; ia <= SSRG_QANG; ia += 64) {
      const double th = (double)(ia - 1) * PI / (double)(SSRG_QANG - 1);
      const double x = k * dw * sin(th * 0.5);
      const double ct = cos(th);
      last = pref * ssrg_shape(x, kappa, beta, gamma, zeta, tg) * (1.0 + ct * ct) * 0.5;
      q += last * sin(th);
    }
    q = wave_sum(q);
    const double qback = __shfl(last, LAST_LANE, 64);
    if (lane == 0) {
      const double qscat = 0.5 * PI * q / (double)SSRG_QANG;
      const double qext = qscat + 3.0 * k * vol * K.im;
      bq[ii] = qext * nd[ii] * dd[ii];
      bq[MIE_MAXBIN + ii] = qscat * nd[ii] * dd[ii];
      bq[2 * MIE_MAXBIN + ii] = qback * nd[ii] * dd[ii];
    }
  }
  // amplitude-matrix sums at the Gauss nodes (rayleigh_gans.f90:733-744)
  if (t < nquad) {
    const double th = acos(tmu[t]);
    const double ct = cos(th), sh = sin(th * 0.5);
    const double k3 = k2 * k;
    double p1 = 0.0, p2 = 0.0, p3 = 0.0;
    for (int ii = 0; ii < nbin; ii++) {
      const double n = nd[ii];
      if (!(n > 0.0)) continue;
      const double dw = (pb.canting(ii) == 0.0) ? dia[ii] * pb.as_ratio(ii) : dia[ii];
      const double vol = mass[ii] / 917.0;
      const double shape =
          ssrg_shape(k * dw * sh, pb.kappa(ii), pb.beta(ii), pb.gamma(ii), pb.zeta(ii), tg);
      const double amp = 3.0 * k3 * vol * sqrt(shape) / (4.0 * PI);
      // s22 = -i amp K, s11 = s22 cos(theta)
      const double s22r = amp * K.im, s22i = -amp * K.re;
      const double s11r = s22r * ct, s11i = s22i * ct;
      const double a11 = s11r * s11r + s11i * s11i, a22 = s22r * s22r + s22i * s22i;
      p1 = p1 + 0.5 * (a11 + a22) * n * dd[ii];
      p2 = p2 + 0.5 * (a11 - a22) * n * dd[ii];
      p3 = p3 + (s11r * s22r + s11i * s22i) * n * dd[ii];
    }
    s_ps[0][t] = p1;
    s_ps[1][t] = p2;
    s_ps[2][t] = p3;
  }
  __syncthreads();
  if (t == 0) {  // bin order, as the reference sums
    double e = 0.0, sc = 0.0, b = 0.0;
    for (int ii = 0; ii < nbin; ii++) {
      e = e + bq[ii];
      sc = sc + bq[MIE_MAXBIN + ii];
      b = b + bq[2 * MIE_MAXBIN + ii];
    }
    sumq[0] = e;
    sumq[1] = sc;
    sumq[2] = b;
  }
  return true;
}

// One (scattering layer, channel) unit: Mie of every hydrometeor of the
// layer and their accumulation into the layer's Legendre series.
// SSRG: the instantiation that also carries the self-similar Rayleigh-Gans
// branch (its registers would otherwise spill in the Mie-only kernel).
template <bool SSRG>
__device__ __forceinline__ void mie_unit(const Dims &d, const Settings &s,
                                         const pam_hydro_desc *hyd, const Hydro4D &h4,
                                         const double *freqs, const Work &w, int itf,
                                         MieLds &L) {
  double *s_u = L.u;
  double(*s_bin)[MIE_WG] = L.bin;
  double *s_wb = L.wb;
  double(*s_ps)[MAX_NQUAD] = L.ps;
  double *s_tab = L.tab;
  double(*s_pw)[6 * 16] = L.pw;
  double(*s_lc)[MAXNLEG + 1] = L.lc;
  double(*s_red)[3] = L.red;
  int(*s_redi)[2] = L.redi;
  double(*s_hm)[2] = L.hm;
  int *s_hnb = L.hnb, *s_hnt = L.hnt;
  short *s_nmb = L.nmb, *s_cb = L.cb, *s_cnp = L.cnp, *s_cnm = L.cnm;
  int &s_nc = L.nc;
  double *s_sum = L.sum;
  int &s_nlh = L.nlh, &s_nan = L.nan_;

  const int t = threadIdx.x;
  const int wv = t >> 6, lane = t & 63, li = lane & 15, lk = lane >> 4;
  const int li_item = itf / d.nf, f = itf % d.nf;
  const int col = w.item_col[li_item], z = w.item_layer[li_item];
  int32_t *st = w.itf_status + itf;
  const size_t oy = (size_t)col * d.ML + z;
  const double freq = freqs[f];
  const double wavelength = C_LIGHT / (freq * 1e9);

  for (int k = t; k < NCOEF * (MAXNLEG + 1); k += MIE_WG) (&s_lc[0][0])[k] = 0.0;
  // per hydrometeor, one lane each: bins, refractive index
  // (scatProperties.f90:227-243) and the series length of the largest bin
  // (mie_spheres.f90:139-158)
  if (t < d.nh) {
    const int nbin = w.dsd_flag[(size_t)li_item * d.nh + t];
    int nterms = 0;
    cd m = cd{0.0, 0.0};
    if (nbin > 0) {
      const double T = w.temp[oy];
      if (hyd[t].liq_ice == 1)
        m = csqrt_(eps_water_tkc(T - sp(273.15), freq));
      else
        m = csqrt_(eps_ice(T, freq));
      const double *dia = w.dsd_d + ((size_t)li_item * d.nh + t) * d.nbmax;
      nterms = mie_nstop(PI * dia[nbin - 1] / wavelength);
    }
    s_hnb[t] = nbin;
    s_hnt[t] = nterms;
    s_hm[t][0] = m.re;
    s_hm[t][1] = m.im;
  }
  __syncthreads();
  double kexttot = 0.0, salbedo = 0.0, backtot = 0.0;
  int nleg_coef = 0;
  long long work_bins = 0, work_series = 0, work_ang = 0, work_proj = 0;

  for (int hh = 0; hh < d.nh; hh++) {
    const int nbin = s_hnb[hh];
    if (nbin <= 0) continue;
    const size_t ob = ((size_t)li_item * d.nh + hh) * d.nbmax;
    const double *dia = w.dsd_d + ob, *dd = w.dsd_dd + ob, *nd = w.dsd_nd + ob,
                 *rho = w.dsd_rho + ob;
    const cd m_ice = cd{s_hm[hh][0], -s_hm[hh][1]};
    const bool is_ice = (hyd[hh].liq_ice == -1);
    const int nterms = s_hnt[hh];
    if (nterms < 1 || nterms > w.atab_nmax || 16 * (2 * ((nterms + 3) & ~3) + 1) > U_CAP ||
        nbin > MIE_MAXBIN) {
      if (t == 0) set_err(st, E_MIE_TERMS);
      return;
    }
    const int nlegen = mie_nlegen(nterms);
    const int nquad = mie_nquad(nterms);
    const double *tmu = w.atab + w.atab_off[nterms], *twt = tmu + nquad, *tpi = twt + nquad,
                 *ttau = tpi + (size_t)nterms * nquad, *tpl = ttau + (size_t)nterms * nquad;
    // the Gauss nodes come in pairs -mu, +mu (the middle node is mu = 0 when
    // nquad is odd): phase B works on the nh2 angles with mu >= 0
    const int nh2 = (nquad + 1) >> 1, j0 = nquad - nh2;
    const int nct = (nh2 + 15) >> 4;
    for (int k = t; k < 3 * MAX_NQUAD; k += MIE_WG) (&s_ps[0][0])[k] = 0.0;
    double sumqe = 0.0, sumqs = 0.0, sumqb = 0.0;
    if (SSRG && hyd[hh].scat_name == PAM_SCAT_SSRG_RT3) {
      __syncthreads();
      double q3[3];
      const pam_hydro_desc &hd = hyd[hh];
      SsrgBins pb{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                  w.dsd_ar[(size_t)li_item * d.nh + hh], hd.canting, hd.rg_kappa, hd.rg_beta,
                  hd.rg_gamma, hd.rg_zeta};
      if (h4.fs_nbin > 0) {  // the full spectrum's per-bin lists
        const size_t b5 = (oy * d.nh + hh) * (size_t)h4.fs_nbin;
        pb.ar = h4.fs_as_ratio + b5;
        pb.cant = h4.fs_canting + b5;
        pb.kap = h4.fs_rg_kappa + b5;
        pb.bet = h4.fs_rg_beta + b5;
        pb.gam = h4.fs_rg_gamma + b5;
        pb.zet = h4.fs_rg_zeta + b5;
      }
      if (!ssrg_sums(hd, pb, nbin, dia, dd, nd, rho, cd{s_hm[hh][0], s_hm[hh][1]}, wavelength,
                     nquad, tmu, s_u, s_ps, q3, &s_nan, st))
        return;
      if (t == 0) {
        sumqe = q3[0];
        sumqs = q3[1];
        sumqb = q3[2];
        work_bins += nbin;
      }
    } else {
    // series length of every bin (mie_scat_utilities.f90:34)
    for (int e = t; e < nbin; e += MIE_WG) {
      int v = 0;
      if (nd[e] > 0.0) {
        const int ns = mie_nstop(PI * dia[e] / wavelength);
        if (ns > nterms) set_err(st, E_MIE_TERMS);
        v = min(ns, nterms);
      }
      s_nmb[e] = (short)v;
    }
    // c_n pi_n, c_n tau_n at the angles mu >= 0 into LDS when they fit
    const bool tab_lds = nterms * nh2 <= MIE_TAB;
    if (tab_lds)
      for (int e = t; e < nterms * nh2; e += MIE_WG) {
        const int n = e / nh2, h = e - n * nh2;
        s_tab[e] = tpi[(size_t)n * nquad + j0 + h];
        s_tab[MIE_TAB + e] = ttau[(size_t)n * nquad + j0 + h];
      }
    __syncthreads();
    // chunk plan (wave 0): consecutive bins (ascending size) as long as their
    // rows of U, with the chunk's own series length, fit the LDS stage.  A
    // window of 64 bins at a time: inclusive prefix max of the series
    // lengths, then the first bin that no longer fits ends the chunk.
    if (wv == 0) {
      int nc = 0, b = 0;
      while (b < nbin && nc < MIE_MAXCHUNK) {
        int nmax = 0, e = b;
        for (;;) {
          const int idx = e + lane;
          int v = max(nmax, idx < nbin ? (int)s_nmb[idx] : 0);
#pragma unroll
          for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(v, o, 64);
            if (lane >= o) v = max(v, u);
          }
          const int npc = max(4, (v + 3) & ~3);
          const bool ok = idx < nbin && idx - b < MIE_WG &&
                          16 * ((idx - b + 8) >> 3) * (2 * npc + 1) <= U_CAP;
          const unsigned long long bad = __ballot(!ok);
          if (bad) {
            const int fl = __ffsll((long long)bad) - 1;
            const int vprev = __shfl(v, fl > 0 ? fl - 1 : 0, 64);
            if (fl > 0) nmax = vprev;
            e += fl;
            break;
          }
          nmax = __shfl(v, 63, 64);
          e += 64;
        }
        if (lane == 0) {
          s_cb[nc] = (short)b;
          s_cnp[nc] = (short)max(4, (nmax + 3) & ~3);
          s_cnm[nc] = (short)nmax;
        }
        nc++;
        b = e;
      }
      if (lane == 0) {
        s_cb[nc] = (short)b;
        s_nc = (b < nbin) ? -1 : nc;
      }
    }
    __syncthreads();
    if (s_nc < 0) {
      if (t == 0) set_err(st, E_MIE_TERMS);
      return;
    }

    for (int c = 0; c < s_nc; c++) {
      const int b0 = s_cb[c], nb = s_cb[c + 1] - b0;
      const int np = s_cnp[c], kp = 2 * np, ldu = kp + 1, nlev = s_cnm[c];
      const int nrt = (nb + 7) >> 3;
      for (int e = t; e < 16 * nrt * ldu; e += MIE_WG) s_u[e] = 0.0;
      __syncthreads();
      // --- A1: threads over bins: the two recurrences of miecalc
      // (mie_scat_utilities.f90:44-72).  Row re/im of bin t: D_n in the
      // a-columns n-1, psi_n / chi_n in the b-columns n-1, psi_0 / chi_0 in
      // the pad column kp.
      if (t < nb) {
        const int ir = b0 + t;
        const int nmie = s_nmb[ir];
        if (nmie > 0) {
          // soft sphere: index of the ice-air mixture at the bin's density
          // (mie_spheres.f90:187-192), decided per bin
          const cd msp = (is_ice && rho[ir] != 917.0) ? eps_mix_air(m_ice, rho[ir]) : m_ice;
          const double x = PI * dia[ir] / wavelength;
          const cd mc = conj(msp);
          const cd y = mc * x;
          const cd yinv = cdiv_m(cd{1.0, 0.0}, y);
          const double xinv = rcp_nr(x);
          double *Ure = s_u + u_row_re(t) * ldu, *Uim = Ure + 4 * ldu;
          // downward log-derivative recursion from n = nmie + 15
          cd dn = cd{0.0, 0.0};
          for (int n = nmie + 15; n >= 2; n--) {
            const cd ny = (double)n * yinv;
            dn = ny - cdiv_m(cd{1.0, 0.0}, dn + ny);
            if (n - 1 <= nmie) {
              Ure[n - 2] = dn.re;
              Uim[n - 2] = dn.im;
            }
          }
          // upward Riccati-Bessel recursion
          double psim = cos(x), psin = sin(x), chim = -sin(x), chin = cos(x);
          Ure[kp] = psin;
          Uim[kp] = chin;
          for (int n = 1; n <= nmie; n++) {
            const double fn = (double)(2 * n - 1) * xinv;
            double tmp = psin;
            psin = fn * psin - psim;
            psim = tmp;
            tmp = chin;
            chin = fn * chin - chim;
            chim = tmp;
            Ure[np + n - 1] = psin;
            Uim[np + n - 1] = chin;
          }
          s_bin[0][t] = mc.re;
          s_bin[1][t] = mc.im;
          s_bin[2][t] = xinv;
        }
        s_wb[t] = nmie > 0 ? nd[ir] * dd[ir] : 0.0;
      } else if (t < 8 * nrt) {
        s_wb[t] = 0.0;
      }
      __syncthreads();
      // --- A2: threads over (bin, n): a_n, b_n (mie_scat_utilities.f90:55-72)
      // over the recurrence values.  Items run in descending n (level-major),
      // and every item reads its inputs before the barrier that precedes the
      // writes, so the b-column n-1 (psi_n, also read by item n+1) is
      // overwritten only after its last reader.
      for (int k0 = 0; k0 < nb * nlev; k0 += MIE_WG) {
        const int k = k0 + t;
        const int bl = k % nb, n = nlev - k / nb;
        const bool ok = k < nb * nlev && n <= s_nmb[b0 + bl];
        cd an = cd{0.0, 0.0}, bn = cd{0.0, 0.0};
        double *Ure = s_u + u_row_re(bl) * ldu, *Uim = Ure + 4 * ldu;
        if (ok) {
          const cd mc = cd{s_bin[0][bl], s_bin[1][bl]};
          const cd mcinv = cdiv_m(cd{1.0, 0.0}, mc);
          const double xinv = s_bin[2][bl];
          const double psin = Ure[np + n - 1], chin = Uim[np + n - 1];
          const double psim = (n == 1) ? Ure[kp] : Ure[np + n - 2];
          const double chim = (n == 1) ? Uim[kp] : Uim[np + n - 2];
          const cd xin = cd{sp(psin), sp(-chin)};
          const cd xim = cd{sp(psim), sp(-chim)};
          const cd Dn = cd{Ure[n - 1], Uim[n - 1]};
          const double nx = (double)n * xinv;
          cd ct = Dn * mcinv + nx;
          an = cdiv_m(ct * psin - psim, ct * xin - xim);
          ct = mc * Dn + nx;
          bn = cdiv_m(ct * psin - psim, ct * xin - xim);
        }
        __syncthreads();
        if (ok) {
          Ure[n - 1] = an.re;
          Uim[n - 1] = an.im;
          Ure[np + n - 1] = bn.re;
          Uim[np + n - 1] = bn.im;
        }
      }
      __syncthreads();
      // --- efficiencies (miecross, mie_scat_utilities.f90:77-116) of bin t,
      // its sums over n in ascending order
      {
        double q0 = 0.0, q1 = 0.0, q2 = 0.0;
        int nbw = 0, nsw = 0;
        if (t < nb && s_nmb[b0 + t] > 0) {
          const int ir = b0 + t, nmie = s_nmb[ir];
          const double *Ure = s_u + u_row_re(t) * ldu, *Uim = Ure + 4 * ldu;
          double s1 = 0.0, s2 = 0.0;
          cd s3 = cd{0.0, 0.0};
          for (int n = 1; n <= nmie; n++) {
            const double wn = (double)(2 * n + 1);
            const double ar = Ure[n - 1], ai = Uim[n - 1], br = Ure[np + n - 1],
                         bi = Uim[np + n - 1];
            s1 = s1 + wn * (ar + br);
            s2 = s2 + wn * ((ar * ar + ai * ai) + (br * br + bi * bi));
            const double ws = (n & 1) ? -wn : wn;
            s3 = s3 + ws * cd{ar - br, ai - bi};
          }
          const double x = PI * dia[ir] / wavelength;
          const double rr = dia[ir] / 2.0;
          const double ar = (rr * rr) * PI;
          const double qe = 2.0 / (x * x) * s1, qs = 2.0 / (x * x) * s2;
          const double qb = ((1.0 / (x * x) * s3) * conj(s3)).re;
          q0 = qe * ar * nd[ir] * dd[ir];
          q1 = qs * ar * nd[ir] * dd[ir];
          q2 = qb * ar * nd[ir] * dd[ir];
          nbw = 1;
          nsw = nmie;
        }
        q0 = wave_sum(q0);
        q1 = wave_sum(q1);
        q2 = wave_sum(q2);
        nbw = wave_sum(nbw);
        nsw = wave_sum(nsw);
        if (lane == 0) {
          s_red[wv][0] = q0;
          s_red[wv][1] = q1;
          s_red[wv][2] = q2;
          s_redi[wv][0] = nbw;
          s_redi[wv][1] = nsw;
        }
      }
      // --- B: S1, S2 of every (bin, angle) on the matrix cores and the
      // DSD-weighted phase-function sums (mieangle, mie_spheres.f90:238-247).
      // With pi_n(-mu) = (-1)^(n-1) pi_n(mu) and tau_n(-mu) = (-1)^n tau_n(mu)
      // the series splits by the parity of n: over the a_n of odd n and the
      // b_n of even n (group 0) against [pi; tau] gives X0 and against
      // [tau; pi] gives Y0, group 1 likewise X1, Y1; then
      //   S1(+mu) = X0 + X1, S1(-mu) = X0 - X1,
      //   S2(+mu) = Y1 + Y0, S2(-mu) = Y1 - Y0,
      // half the products of evaluating both angles.  With four or more
      // tiles of angles (mu >= 0) wave wv owns the tiles ct = wv, wv + 4, ...
      // with every row tile; with fewer, wpc = 4 / nct waves share a tile,
      // each over every wpc-th row tile, and their sums are added in wave
      // order after the barrier.
      const int wpc = nct >= 4 ? 1 : 4 / nct;
      for (int ct = (wpc > 1 ? wv / wpc : wv); ct < nct; ct += (wpc > 1 ? nct : 4)) {
        const int sub = wpc > 1 ? wv % wpc : 0;
        const int hix = 16 * ct + li;
        const bool ang_ok = hix < nh2;
        const int ang = j0 + hix, angm = nquad - 1 - ang;
        const int nsk = np >> 2;  // steps per group
        double p1 = 0.0, p2 = 0.0, p3 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
        for (int rt = sub; rt < nrt; rt += wpc) {
          d4m_t x0 = {0.0, 0.0, 0.0, 0.0}, y0 = x0, x1 = x0, y1 = x0;
          const double *urow = s_u + (16 * rt + li) * ldu;
          for (int sk = 0; sk < nsk; sk++) {
#pragma unroll
            for (int g = 0; g < 2; g++) {
              int cu, n;
              bool ap;
              mie_kmap(g, 4 * sk + lk, np, cu, n, ap);
              double pv = 0.0, tv = 0.0;
              if (ang_ok && n <= nterms) {
                if (tab_lds) {
                  pv = s_tab[(n - 1) * nh2 + hix];
                  tv = s_tab[MIE_TAB + (n - 1) * nh2 + hix];
                } else {
                  pv = tpi[(size_t)(n - 1) * nquad + ang];
                  tv = ttau[(size_t)(n - 1) * nquad + ang];
                }
              }
              const double a = urow[cu];
              if (g == 0) {
                x0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, ap ? pv : tv, x0, 0, 0, 0);
                y0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, ap ? tv : pv, y0, 0, 0, 0);
              } else {
                x1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, ap ? pv : tv, x1, 0, 0, 0);
                y1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, ap ? tv : pv, y1, 0, 0, 0);
              }
            }
          }
#pragma unroll
          for (int h2 = 0; h2 < 2; h2++) {
            const double wgt = s_wb[8 * rt + 4 * h2 + lk];
            const int re = 2 * h2, im = 2 * h2 + 1;
            // +mu
            double s1r = x0[re] + x1[re], s1i = x0[im] + x1[im];
            double s2r = y1[re] + y0[re], s2i = y1[im] + y0[im];
            double a1 = s1r * s1r + s1i * s1i, a2 = s2r * s2r + s2i * s2i;
            p1 += (0.5 * (a2 + a1)) * wgt;
            p2 += (0.5 * (a2 - a1)) * wgt;
            p3 += (s2r * s1r + s2i * s1i) * wgt;
            // -mu
            s1r = x0[re] - x1[re];
            s1i = x0[im] - x1[im];
            s2r = y1[re] - y0[re];
            s2i = y1[im] - y0[im];
            a1 = s1r * s1r + s1i * s1i;
            a2 = s2r * s2r + s2i * s2i;
            m1 += (0.5 * (a2 + a1)) * wgt;
            m2 += (0.5 * (a2 - a1)) * wgt;
            m3 += (s2r * s1r + s2i * s1i) * wgt;
          }
        }
        // the four lane groups hold different bins of the same angle
        p1 += __shfl_xor(p1, 16, 64);
        p1 += __shfl_xor(p1, 32, 64);
        p2 += __shfl_xor(p2, 16, 64);
        p2 += __shfl_xor(p2, 32, 64);
        p3 += __shfl_xor(p3, 16, 64);
        p3 += __shfl_xor(p3, 32, 64);
        m1 += __shfl_xor(m1, 16, 64);
        m1 += __shfl_xor(m1, 32, 64);
        m2 += __shfl_xor(m2, 16, 64);
        m2 += __shfl_xor(m2, 32, 64);
        m3 += __shfl_xor(m3, 16, 64);
        m3 += __shfl_xor(m3, 32, 64);
        if (lk == 0 && ang_ok) {
          if (wpc == 1) {  // this wave alone owns the pair
            s_ps[0][ang] += p1;
            s_ps[1][ang] += p2;
            s_ps[2][ang] += p3;
            if (angm != ang) {
              s_ps[0][angm] += m1;
              s_ps[1][angm] += m2;
              s_ps[2][angm] += m3;
            }
          } else {
            double *pw = s_pw[wv];
            pw[li] = p1;
            pw[16 + li] = p2;
            pw[32 + li] = p3;
            pw[48 + li] = m1;
            pw[64 + li] = m2;
            pw[80 + li] = m3;
          }
        }
      }
      __syncthreads();
      if (wpc > 1 && t < 16 * nct) {  // the shared tiles, waves in order
        const int ct = t >> 4, l16 = t & 15, hix = t;
        if (hix < nh2) {
          const int ang = j0 + hix, angm = nquad - 1 - ang;
          double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
          for (int q = 0; q < wpc; q++) {
            const double *pw = s_pw[ct * wpc + q];
#pragma unroll
            for (int k = 0; k < 6; k++) v[k] += pw[16 * k + l16];
          }
          s_ps[0][ang] += v[0];
          s_ps[1][ang] += v[1];
          s_ps[2][ang] += v[2];
          if (angm != ang) {
            s_ps[0][angm] += v[3];
            s_ps[1][angm] += v[4];
            s_ps[2][angm] += v[5];
          }
        }
      }
      if (t == 0) {  // the chunk's efficiency sums, waves in order
        for (int q = 0; q < 4; q++) {
          sumqe = sumqe + s_red[q][0];
          sumqs = sumqs + s_red[q][1];
          sumqb = sumqb + s_red[q][2];
          work_bins += s_redi[q][0];
          work_series += s_redi[q][1];
          work_ang += (long long)nquad * s_redi[q][1];
        }
      }
    }
    }  // Mie spheres
    if (t == 0) {
      s_sum[0] = sumqe;
      s_sum[1] = sumqs;
      s_sum[2] = sumqb;
      s_nlh = 0;
      s_nan = 0;
      work_proj += (long long)nquad * (nlegen + 1);
    }
    __syncthreads();
    sumqe = s_sum[0];
    sumqs = s_sum[1];
    sumqb = s_sum[2];
    const double extinction = s.lhyd_abs ? sumqe : sumqe - sumqs;
    const double scatter = sumqs;
    const double albedo = scatter / extinction;
    if (!(extinction > 0.0) || !(scatter > 0.0) || !(sumqb > 0.0)) {
      if (t == 0) set_err(st, E_MIE_RESULT);
      return;
    }
    // weighted phase sums per angle, normalised (mie_spheres.f90:272-280),
    // into the U stage
    double *s_q = s_u;
    const double tmp = wavelength * wavelength / (PI * scatter);
    if (t < nquad) {
      const double wt = twt[t];
      s_q[t] = tmp * s_ps[0][t] * wt;
      s_q[MAX_NQUAD + t] = tmp * s_ps[1][t] * wt;
      s_q[2 * MAX_NQUAD + t] = tmp * s_ps[2][t] * wt;
    }
    __syncthreads();
    // --- C: Legendre projection (mie_spheres.f90:272-314), thread per
    // order, over the pairs of angles: P_l(-mu) = (-1)^l P_l(mu)
    double g1 = 0.0, g2 = 0.0, g3 = 0.0;
    if (t <= nlegen) {
      double c1 = 0.0, c2 = 0.0, c3 = 0.0;
      const double sg = (t & 1) ? -1.0 : 1.0;
#pragma unroll 4
      for (int h = 0; h < nh2; h++) {
        const int k = j0 + h, km = nquad - 1 - k;
        const double pl = tpl[(size_t)k * (nlegen + 1) + t];
        double q1 = s_q[k], q2 = s_q[MAX_NQUAD + k], q3 = s_q[2 * MAX_NQUAD + k];
        if (km != k) {
          q1 += sg * s_q[km];
          q2 += sg * s_q[MAX_NQUAD + km];
          q3 += sg * s_q[2 * MAX_NQUAD + km];
        }
        c1 += q1 * pl;
        c2 += q2 * pl;
        c3 += q3 * pl;
      }
      const double hl = (double)((float)(2 * t + 1) / 2.0f);
      g1 = hl * c1;
      g2 = hl * c2;
      g3 = hl * c3;
      if (g1 > sp(1.0e-7)) atomicMax(&s_nlh, t);
      if (isnan(g1) || isnan(c2) || isnan(c3)) s_nan = 1;
    }
    if (t == 0 && s.fault == FAULT_NAN_COEF && col == s.fault_col && f == s.fault_ch) s_nan = 1;
    __syncthreads();
    if (s_nan) {
      if (t == 0) set_err(st, E_NAN_COEF);
      return;
    }
    // accumulate over hydrometeors (scatProperties.f90:920-946)
    const int nl_h = s_nlh;
    if (nl_h > nleg_coef) nleg_coef = nl_h;
    kexttot = kexttot + extinction;
    backtot = backtot + sumqb;  // rt_back (scatProperties.f90:743-757, 922)
    salbedo = (kexttot == 0.0) ? 0.0 : salbedo + (albedo * extinction);
    if (t < nleg_coef && t <= nlegen) {
      s_lc[0][t] = s_lc[0][t] + (g1 * albedo * extinction);
      s_lc[1][t] = s_lc[1][t] + (g2 * albedo * extinction);
      s_lc[2][t] = s_lc[2][t] + (g3 * albedo * extinction);
    }
    __syncthreads();
    if (kexttot < 0.0 || isnan(kexttot)) {
      if (t == 0) set_err(st, E_KEXT);
      return;
    }
    if (nleg_coef > MAXNLEG - 1) {
      if (t == 0) set_err(st, E_NLEG);
      return;
    }
  }
  // finalize_rt3_scatProperties and output
  salbedo = salbedo / kexttot;
  const double norm = salbedo * kexttot;
  double *out = w.coef + (size_t)itf * NCOEF * MAXNLEG;
  for (int l = t; l <= nleg_coef && l < MAXNLEG; l += MIE_WG) {
    out[l] = s_lc[0][l] / norm;
    out[MAXNLEG + l] = s_lc[1][l] / norm;
    out[2 * MAXNLEG + l] = s_lc[2][l] / norm;
  }
  if (t == 0) {
    w.kext[itf] = kexttot;
    w.back[itf] = backtot;
    w.salb[itf] = salbedo;
    w.nleg[itf] = nleg_coef;
    atomicAdd(&w.counters[6], (unsigned long long)work_bins);
    atomicAdd(&w.counters[7], (unsigned long long)work_series);
    atomicAdd(&w.counters[12], (unsigned long long)work_ang);
    atomicAdd(&w.counters[13], (unsigned long long)work_proj);
  }
}

// ---------------------------------------------------------------------
// Mie spheres, coefficient-first form (the kernel for calls without
// self-similar Rayleigh-Gans hydrometeors).  One 128-thread workgroup per
// (scattering layer, channel) unit, persistent, with a slot of global
// scratch per workgroup.  Per hydrometeor:
//   A. threads over all size bins at once: the D_n recursion, the upward
//      psi_n/chi_n recursion with a_n, b_n (miecalc) and the efficiencies
//      (miecross) in one serial pass per bin; a_n, b_n go to the slot as
//      the U matrix of the S1/S2 product (column-major, row pairs re/im per
//      bin as in u_row_re), D_n passes through the a-columns;
//   B. waves over (angle tile, row tile): S1/S2 on the FP64 matrix cores
//      with U read from the slot (L2-resident, just written by this
//      workgroup) and series terms beyond a bin's own length read as zero;
//      the DSD-weighted phase sums into LDS;
//   C. threads over Legendre orders, kept in registers across hydrometeors.
// No stage is chunked over bins, so a hydrometeor costs one pass of each
// phase and four barriers, and the small LDS footprint keeps several units
// resident per compute unit.
// ---------------------------------------------------------------------
constexpr int MIE2_WG = 128;
constexpr int MIE2_NW = MIE2_WG / 64;
constexpr int MIE2_TAB = 768;          // LDS stage of c_n pi_n, c_n tau_n (n x angles)
// extinction [1/m] below which the mixed path skips a size bin: over a 20 km
// column a layer of such bins changes the optical depth by < 1e-8 per bin
constexpr double MIE2_CULL_EXT = 1e-13;
// a bin of number density nd [1/m^4], width dd and diameter dia [m] whose
// extinction is bounded below MIE2_CULL_EXT (cross section <= 5 pi r^2)
__device__ __forceinline__ bool mie2_negligible(double nd, double dd, double dia) {
  const double r = 0.5 * dia;
  return nd * dd * (5.0 * PI * r * r) < MIE2_CULL_EXT;
}
constexpr int MIE2_NL = (MAXNLEG + MIE2_WG - 1) / MIE2_WG;  // Legendre orders per thread

// Rows of the U slot.  FP64 (16x16x4 f64 MFMA, lane group k holds result
// rows k + 4 r): per 16-row tile [re b0..b3, im b0..b3, re b4..b7, im
// b4..b7].  FP32 (16x16x4 f32 MFMA, lane group k holds rows 4 k + r): per
// tile [re b0, im b0, re b1, im b1, ...].  Either way a lane's four results
// are the real and imaginary parts of two bins at one angle.
template <typename R> struct Mie2Rows;
template <> struct Mie2Rows<double> {
  typedef d4m_t acc_t;
  static constexpr int im_off = 4;
  static __device__ __forceinline__ int row_re(int b) { return 16 * (b >> 3) + u_row_re(b & 7); }
  static __device__ __forceinline__ int tile_bin(int r) { return (r & 3) + 4 * (r >> 3); }
  // bin (within the tile) of result registers 2 h, 2 h + 1 of lane group k
  static __device__ __forceinline__ int out_bin(int k, int h) { return 4 * h + k; }
  static __device__ __forceinline__ acc_t mfma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct Mie2Rows<float> {
  typedef f4m_t acc_t;
  static constexpr int im_off = 1;
  static __device__ __forceinline__ int row_re(int b) { return 16 * (b >> 3) + 2 * (b & 7); }
  static __device__ __forceinline__ int tile_bin(int r) { return r >> 1; }
  static __device__ __forceinline__ int out_bin(int k, int h) { return 2 * k + h; }
  static __device__ __forceinline__ acc_t mfma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// miecalc + miecross for one bin (mie_scat_utilities.f90:8-116), FP64: the
// D_n recursion, then psi_n, chi_n, a_n, b_n and the efficiency sums in
// ascending n.  a_n, b_n go to the slot rows at ure (type R); D_n is parked
// in the a-columns between the two passes.
template <typename R>
__device__ __forceinline__ void mie2_series_f64(double x, cd msp, int nmie, int nph, int ld, R *ure,
                                                double &s1, double &s2, cd &s3) {
  const cd mc = conj(msp);
  const cd y = mc * x;
  const cd yinv = cdiv_m(cd{1.0, 0.0}, y);
  const double xinv = rcp_nr(x);
  const cd mcinv = cdiv_m(cd{1.0, 0.0}, mc);
  R *uim = ure + Mie2Rows<R>::im_off;
  // downward log-derivative recursion from n = nmie + 15; D_n (n <= nmie)
  // parked in the a-column n-1
  cd dn = cd{0.0, 0.0};
  for (int n = nmie + 15; n >= 2; n--) {
    const cd ny = (double)n * yinv;
    dn = ny - cdiv_m(cd{1.0, 0.0}, dn + ny);
    if (n - 1 <= nmie) {
      ure[(size_t)(n - 2) * ld] = (R)dn.re;
      uim[(size_t)(n - 2) * ld] = (R)dn.im;
    }
  }
  // upward Riccati-Bessel recursion, a_n, b_n and the efficiency sums in
  // ascending n
  double psim = cos(x), psin = sin(x), chim = -sin(x), chin = cos(x);
  s1 = 0.0;
  s2 = 0.0;
  s3 = cd{0.0, 0.0};
  // D_n is read back four terms ahead of its use (dq[] rotates), the
  // loads issued before the stores of the terms between
  cd dq[4];
#pragma unroll
  for (int q = 0; q < 4; q++)
    dq[q] = (q < nmie) ? cd{(double)ure[(size_t)q * ld], (double)uim[(size_t)q * ld]} : cd{0.0, 0.0};
  for (int n0 = 1; n0 <= nmie; n0 += 4)
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int n = n0 + q;
    if (n > nmie) break;
    const cd Dn = dq[q];
    if (n + 4 <= nmie) dq[q] = cd{(double)ure[(size_t)(n + 3) * ld], (double)uim[(size_t)(n + 3) * ld]};
    const double fn = (double)(2 * n - 1) * xinv;
    double tmp = psin;
    psin = fn * psin - psim;
    psim = tmp;
    tmp = chin;
    chin = fn * chin - chim;
    chim = tmp;
    const cd xin = cd{sp(psin), sp(-chin)};
    const cd xim = cd{sp(psim), sp(-chim)};
    const double nx = (double)n * xinv;
    cd ct = Dn * mcinv + nx;
    const cd an = cdiv_m(ct * psin - psim, ct * xin - xim);
    ct = mc * Dn + nx;
    const cd bn = cdiv_m(ct * psin - psim, ct * xin - xim);
    ure[(size_t)(n - 1) * ld] = (R)an.re;
    uim[(size_t)(n - 1) * ld] = (R)an.im;
    ure[(size_t)(nph + n - 1) * ld] = (R)bn.re;
    uim[(size_t)(nph + n - 1) * ld] = (R)bn.im;
    const double wn = (double)(2 * n + 1);
    s1 = s1 + wn * (an.re + bn.re);
    s2 = s2 + wn * ((an.re * an.re + an.im * an.im) + (bn.re * bn.re + bn.im * bn.im));
    const double ws = (n & 1) ? -wn : wn;
    s3 = s3 + ws * cd{an.re - bn.re, an.im - bn.im};
  }
}

// Single-precision complex arithmetic of the mixed-precision Mie series.
struct cf {
  float re, im;
};
__device__ __forceinline__ cf operator+(cf a, cf b) { return cf{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cf operator-(cf a, cf b) { return cf{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cf operator*(cf a, cf b) {
  return cf{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cf operator*(float s, cf a) { return cf{s * a.re, s * a.im}; }
__device__ __forceinline__ cf operator*(cf a, float s) { return cf{a.re * s, a.im * s}; }
__device__ __forceinline__ cf operator+(cf a, float s) { return cf{a.re + s, a.im}; }
__device__ __forceinline__ cf operator-(cf a, float s) { return cf{a.re - s, a.im}; }
// 1/x from the hardware reciprocal and one Newton step
__device__ __forceinline__ float rcp_nrf(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return fmaf(r, fmaf(-x, r, 1.f), r);
}
__device__ __forceinline__ cf cdiv_f(cf a, cf b) {
  const float r = rcp_nrf(b.re * b.re + b.im * b.im);
  return cf{(a.re * b.re + a.im * b.im) * r, (a.im * b.re - a.re * b.im) * r};
}

// The mixed-precision series of one bin: the complex work (the D_n
// recursion, a_n, b_n) in FP32, the real Riccati-Bessel recursion in FP64
// and the efficiency sums accumulated in FP64.
//
// psi_n and chi_n stay in double because the upward recursion cancels at
// small x: psi_1 = sin x / x - cos x is O(x^2) from O(1) terms, so an FP32
// psi_1 carries a relative error ~eps / x^2 (1.5e-2 at x = 2e-3), which
// reaches a_1 through its numerator.  With psi_n, chi_n exact to FP64 and
// rounded once, the remaining FP32 cancellation is in the numerator of b_1,
// (m D_1 + 1/x) psi_1 - psi_0 = O(x^3) from O(x) terms: b_1 then has a
// relative error ~eps / x^2, but |b_1| ~ x^2 |a_1|, so its absolute error
// stays ~eps |a_1| in every sum it enters.  Bins with x < MIE2_X32 (where
// chi_n^2 approaches the FP32 range) take the FP64 series.
constexpr double MIE2_X32 = 1.0e-3;
__device__ __forceinline__ void mie2_series_f32(double x, cd msp, int nmie, int nph, int ld,
                                                float *ure, double &s1, double &s2, cd &s3) {
  float *uim = ure + Mie2Rows<float>::im_off;
  const cf mc = cf{(float)msp.re, -(float)msp.im};
  const float xf = (float)x;
  const cf y = mc * xf;
  const cf yinv = cdiv_f(cf{1.f, 0.f}, y);
  const double xinv = rcp_nr(x);
  const cf mcinv = cdiv_f(cf{1.f, 0.f}, mc);
  cf dn = cf{0.f, 0.f};
  for (int n = nmie + 15; n >= 2; n--) {
    const cf ny = (float)n * yinv;
    dn = ny - cdiv_f(cf{1.f, 0.f}, dn + ny);
    if (n - 1 <= nmie) {
      ure[(size_t)(n - 2) * ld] = dn.re;
      uim[(size_t)(n - 2) * ld] = dn.im;
    }
  }
  double psim = cos(x), psin = sin(x), chim = -sin(x), chin = cos(x);
  s1 = 0.0;
  s2 = 0.0;
  s3 = cd{0.0, 0.0};
  cf dq[4];
#pragma unroll
  for (int q = 0; q < 4; q++)
    dq[q] = (q < nmie) ? cf{ure[(size_t)q * ld], uim[(size_t)q * ld]} : cf{0.f, 0.f};
  for (int n0 = 1; n0 <= nmie; n0 += 4)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int n = n0 + q;
      if (n > nmie) break;
      const cf Dn = dq[q];
      if (n + 4 <= nmie) dq[q] = cf{ure[(size_t)(n + 3) * ld], uim[(size_t)(n + 3) * ld]};
      const double fn = (double)(2 * n - 1) * xinv;
      double tmp = psin;
      psin = fn * psin - psim;
      psim = tmp;
      tmp = chin;
      chin = fn * chin - chim;
      chim = tmp;
      const cf xin = cf{(float)psin, (float)(-chin)};
      const cf xim = cf{(float)psim, (float)(-chim)};
      const float pn = (float)psin, pm = (float)psim;
      const float nx = (float)((double)n * xinv);
      cf ct = Dn * mcinv + nx;
      const cf an = cdiv_f(ct * pn - cf{pm, 0.f}, ct * xin - xim);
      ct = mc * Dn + nx;
      const cf bn = cdiv_f(ct * pn - cf{pm, 0.f}, ct * xin - xim);
      ure[(size_t)(n - 1) * ld] = an.re;
      uim[(size_t)(n - 1) * ld] = an.im;
      ure[(size_t)(nph + n - 1) * ld] = bn.re;
      uim[(size_t)(nph + n - 1) * ld] = bn.im;
      const double wn = (double)(2 * n + 1);
      const double ar = an.re, ai = an.im, br = bn.re, bi = bn.im;
      s1 = s1 + wn * (ar + br);
      s2 = s2 + wn * ((ar * ar + ai * ai) + (br * br + bi * bi));
      const double ws = (n & 1) ? -wn : wn;
      s3 = s3 + ws * cd{ar - br, ai - bi};
    }
}

// R = float: 19,748 B, within the 20 KB of eight workgroups per compute unit
template <typename R>
struct Mie2Lds {
  double ps[3][MAX_NQUAD];     // weighted phase sums per angle
  R tab[2 * MIE2_TAB];         // angle table stage
  short nmb[MIE_MAXBIN];       // per bin: series length (0: skipped)
  R pw[MIE2_NW][6 * 16];       // per-wave phase sums of shared angle tiles
  double red[MIE2_NW][3];
  int redi[MIE2_NW][2];
  double hm[MAX_HYDRO][2];
  int hnb[MAX_HYDRO], hnt[MAX_HYDRO];
  // FP32 instance only (the FP64 one is at its LDS limit of six workgroups per
  // CU): the largest bin kept (-1: none skipped), and whether it is searched
  static constexpr int NHK = sizeof(R) == 4 ? MAX_HYDRO : 1;
  int hkeep[NHK], hscan[NHK];
  double lc[NCOEF][MAXNLEG + 1];  // Legendre sums over hydrometeors (thread l owns order l)
  int nlh, nan_;
  int slot;
};

// R = double: the FP64 kernel.  R = float: the mixed-precision kernel (FP32
// series and S1/S2 products with the slot, the angle tables and the MFMA in
// FP32; the efficiency sums, the phase sums and the Legendre projection in
// FP64).
template <typename R>
__device__ __forceinline__ void mie2_unit(const Dims &d, const Settings &s,
                                          const pam_hydro_desc *hyd, const double *freqs,
                                          const Work &w, int itf, R *U, Mie2Lds<R> &L) {
  constexpr bool F32 = sizeof(R) == 4;
  typedef Mie2Rows<R> M2;
  typedef R PA;  // per-lane phase-sum accumulators
  typedef typename M2::acc_t acc_t;
  // the angle-table stage holds TABN entries of each of c_n pi_n, c_n tau_n
  constexpr int TABN = MIE2_TAB;
  R *const tab = L.tab;
  const int t = threadIdx.x;
  const int wv = t >> 6, lane = t & 63, li = lane & 15, lk = lane >> 4;
  const int li_item = itf / d.nf, f = itf % d.nf;
  const int col = w.item_col[li_item], z = w.item_layer[li_item];
  int32_t *st = w.itf_status + itf;
  const double freq = freqs[f];
  const double wavelength = C_LIGHT / (freq * 1e9);
  const int ld = w.mie_ld;  // rows of the U slot (16 per 8 bins)

  // per hydrometeor, one thread each: bins, refractive index
  // (scatProperties.f90:227-243), series length of the largest bin
  // (mie_spheres.f90:139-158)
  // (the indices of water and ice come from k_mie_index)
  if (t < d.nh) {
    const int nbin = w.dsd_flag[(size_t)li_item * d.nh + t];
    int nterms = 0;
    cd m = cd{0.0, 0.0};
    if (nbin > 0) {
      const double *mi = w.mie_m + (size_t)itf * 4 + (hyd[t].liq_ice == 1 ? 0 : 2);
      m = cd{mi[0], mi[1]};
      const double *dia = w.dsd_d + ((size_t)li_item * d.nh + t) * d.nbmax;
      nterms = mie_nstop(PI * dia[nbin - 1] / wavelength);
    }
    L.hnb[t] = nbin;
    L.hnt[t] = nterms;
    // culling (below): -2 marks a hydrometeor whose largest bin is negligible
    // and whose largest kept bin is searched; -1: no bin skipped
    if constexpr (F32) {
      int keep = -1;
      if (!s.active && nbin > 0) {
        const size_t o = ((size_t)li_item * d.nh + t) * d.nbmax + nbin - 1;
        keep = mie2_negligible(w.dsd_nd[o], w.dsd_dd[o], w.dsd_d[o]) ? -2 : nbin - 1;
      }
      L.hkeep[t] = keep;
      L.hscan[t] = keep == -2;
    }
    L.hm[t][0] = m.re;
    L.hm[t][1] = m.im;
  }
  for (int k = t; k < NCOEF * (MAXNLEG + 1); k += MIE2_WG) (&L.lc[0][0])[k] = 0.0;
  __syncthreads();
  // The mixed path of a passive run skips the bins whose extinction is
  // bounded below MIE2_CULL_EXT (their particles' extinction cross section
  // taken as at most 5 pi r^2): the large bins far in an exponential tail.
  // A hydrometeor's series length, and with it its Gauss quadrature and
  // Legendre orders (mie_spheres.f90:139-158), then follow its largest kept
  // bin.  A hydrometeor none of whose bins would be kept keeps them all.
  bool cull = false, scan = false;
  if constexpr (F32) {
    cull = !s.active;
    for (int hh = 0; hh < d.nh; hh++) scan = scan || L.hscan[hh];
  }
  if constexpr (F32) if (cull && scan) {
    for (int e = t; e < d.nh * d.nbmax; e += MIE2_WG) {
      const int hh = e / d.nbmax, ir = e - hh * d.nbmax;
      if (L.hscan[hh] && ir < L.hnb[hh]) {
        const size_t o = ((size_t)li_item * d.nh + hh) * d.nbmax + ir;
        if (!mie2_negligible(w.dsd_nd[o], w.dsd_dd[o], w.dsd_d[o])) atomicMax(&L.hkeep[hh], ir);
      }
    }
    __syncthreads();
    if (t < d.nh) {
      if (L.hkeep[t] >= 0)
        L.hnt[t] = mie_nstop(PI * w.dsd_d[((size_t)li_item * d.nh + t) * d.nbmax + L.hkeep[t]] /
                             wavelength);
      else
        L.hkeep[t] = -1;  // no bin kept: none skipped
    }
    __syncthreads();
  }
  double kexttot = 0.0, salbedo = 0.0, backtot = 0.0;
  int nleg_coef = 0;
  long long work_bins = 0, work_series = 0, work_ang = 0, work_proj = 0;

  for (int hh = 0; hh < d.nh; hh++) {
    const int nbin = L.hnb[hh];
    if (nbin <= 0) continue;
    const size_t ob = ((size_t)li_item * d.nh + hh) * d.nbmax;
    const double *dia = w.dsd_d + ob, *dd = w.dsd_dd + ob, *nd = w.dsd_nd + ob,
                 *rho = w.dsd_rho + ob;
    const cd m_ice = cd{L.hm[hh][0], -L.hm[hh][1]};
    const bool is_ice = (hyd[hh].liq_ice == -1);
    const int nterms = L.hnt[hh];
    bool hcull = false;
    if constexpr (F32) hcull = cull && L.hkeep[hh] >= 0;
    if (nterms < 1 || nterms > w.atab_nmax || nterms > MAX_NTERMS || nbin > MIE_MAXBIN ||
        16 * ((nbin + 7) >> 3) > ld) {
      if (t == 0) set_err(st, E_MIE_TERMS);
      return;
    }
    const int nph = max(4, (nterms + 3) & ~3);  // b-columns start here
    const int nlegen = mie_nlegen(nterms);
    const int nquad = mie_nquad(nterms);
    const double *tmu = w.atab + w.atab_off[nterms], *twt = tmu + nquad,
                 *tpl = twt + nquad + 2 * (size_t)nterms * nquad;
    const R *tpi = (F32 ? (const R *)w.atab_f : (const R *)w.atab) + w.atab_off[nterms] + 2 * nquad;
    const R *ttau = tpi + (size_t)nterms * nquad;
    const int nh2 = (nquad + 1) >> 1, j0 = nquad - nh2;
    const int nct = (nh2 + 15) >> 4;
    const int nrt = (nbin + 7) >> 3;
    for (int k = t; k < 3 * MAX_NQUAD; k += MIE2_WG) (&L.ps[0][0])[k] = 0.0;
    // c_n pi_n, c_n tau_n at the angles mu >= 0 into LDS when they fit
    // the whole slice at mu >= 0 when it fits; else, for series up to
    // TABN / 16 terms, one 16-angle tile of it per pass of B
    const bool tab_lds = nterms * nh2 <= TABN;
    const bool tab_tile = !tab_lds && nterms * 16 <= TABN;
    auto stage_tile = [&](int ct) {
      for (int e = t; e < nterms * 16; e += MIE2_WG) {
        const int n = e >> 4, hix = 16 * ct + (e & 15);
        const bool ok = hix < nh2;
        tab[e] = ok ? tpi[(size_t)n * nquad + j0 + hix] : (R)0;
        tab[TABN + e] = ok ? ttau[(size_t)n * nquad + j0 + hix] : (R)0;
      }
    };
    if (tab_lds)
      for (int e = t; e < nterms * nh2; e += MIE2_WG) {
        const int n = e / nh2, h = e - n * nh2;
        tab[e] = tpi[(size_t)n * nquad + j0 + h];
        tab[TABN + e] = ttau[(size_t)n * nquad + j0 + h];
      }
    else if (tab_tile)
      stage_tile(0);

    // --- A: one thread per bin, the whole series (mie_scat_utilities.f90:8-116)
    double q0 = 0.0, q1 = 0.0, q2 = 0.0;
    int nbw = 0, nsw = 0;
    for (int ir = t; ir < nbin; ir += MIE2_WG) {
      int nmie = 0;
      if (nd[ir] > 0.0 && !(hcull && mie2_negligible(nd[ir], dd[ir], dia[ir]))) {
        const int ns = mie_nstop(PI * dia[ir] / wavelength);
        if (ns > nterms) set_err(st, E_MIE_TERMS);
        nmie = min(ns, nterms);
      }
      L.nmb[ir] = (short)nmie;
      if (nmie <= 0) continue;
      // soft sphere: index of the ice-air mixture at the bin's density
      // (mie_spheres.f90:187-192), decided per bin
      const cd msp = (is_ice && rho[ir] != 917.0) ? eps_mix_air(m_ice, rho[ir]) : m_ice;
      const double x = PI * dia[ir] / wavelength;
      double s1, s2;
      cd s3;
      if constexpr (F32) {
        if (x >= MIE2_X32)
          mie2_series_f32(x, msp, nmie, nph, ld, U + M2::row_re(ir), s1, s2, s3);
        else
          mie2_series_f64(x, msp, nmie, nph, ld, U + M2::row_re(ir), s1, s2, s3);
      } else {
        mie2_series_f64(x, msp, nmie, nph, ld, U + M2::row_re(ir), s1, s2, s3);
      }
      const double rr = dia[ir] / 2.0;
      const double ar = (rr * rr) * PI;
      const double qe = 2.0 / (x * x) * s1, qs = 2.0 / (x * x) * s2;
      const double qb = ((1.0 / (x * x) * s3) * conj(s3)).re;
      q0 += qe * ar * nd[ir] * dd[ir];
      q1 += qs * ar * nd[ir] * dd[ir];
      q2 += qb * ar * nd[ir] * dd[ir];
      nbw += 1;
      nsw += nmie;
    }
    q0 = wave_sum(q0);
    q1 = wave_sum(q1);
    q2 = wave_sum(q2);
    nbw = wave_sum(nbw);
    nsw = wave_sum(nsw);
    if (lane == 0) {
      L.red[wv][0] = q0;
      L.red[wv][1] = q1;
      L.red[wv][2] = q2;
      L.redi[wv][0] = nbw;
      L.redi[wv][1] = nsw;
    }
    __syncthreads();  // U, the series lengths, weights and table stage are ready

    // --- B: S1, S2 of every (bin, angle) on the matrix cores and the
    // DSD-weighted phase-function sums (mieangle, mie_spheres.f90:238-247),
    // parity-split over mu >= 0 as in the chunked kernel.  Waves own angle
    // tiles; with fewer tiles than waves, wpc waves share a tile, each over
    // every wpc-th row tile.  A row tile's own longest series sets its
    // number of k-steps.  With the table staged one angle tile at a time,
    // the waves take the tiles together, one pass each.
    const int npass = tab_tile ? nct : 1;
    for (int pass = 0; pass < npass; pass++) {
    if (pass > 0) {
      stage_tile(pass);
      __syncthreads();
    }
    const int wpc = tab_tile ? MIE2_NW : (nct >= MIE2_NW ? 1 : MIE2_NW / nct);
    const int ct0 = tab_tile ? pass : (wpc > 1 ? wv / wpc : wv);
    const int ct1 = tab_tile ? pass + 1 : nct;
    for (int ctw = ct0; ctw < ct1; ctw += (wpc > 1 ? nct : MIE2_NW)) {
      const int sub = tab_tile ? wv : (wpc > 1 ? wv % wpc : 0);
      const int hix = 16 * ctw + li;
      const bool ang_ok = hix < nh2;
      const int ang = j0 + hix, angm = nquad - 1 - ang;
      // the phase sums of a lane (FP32 in the mixed kernel: at most a few
      // tens of positive-weighted terms per lane before the FP64 sums)
      PA p1 = 0.0, p2 = 0.0, p3 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
      for (int rt = sub; rt < nrt; rt += wpc) {
        // longest series of the tile's 8 bins, and of this lane's row
        int nmt = 0;
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const int b = 8 * rt + q;
          nmt = max(nmt, b < nbin ? (int)L.nmb[b] : 0);
        }
        if (nmt == 0) continue;  // no bin of the tile scatters: its terms are zero
        const int brow = 8 * rt + M2::tile_bin(li);
        const int nrow = brow < nbin ? (int)L.nmb[brow] : 0;
        const int np = max(4, (nmt + 3) & ~3);
        const int nsk = np >> 2;
        acc_t x0 = {(R)0, (R)0, (R)0, (R)0}, y0 = x0, x1 = x0, y1 = x0;
        // k-step sk of parity group g: the lanes lk < 2 take a_n, the lanes
        // lk >= 2 take b_n, each at its own n = 4 sk + c_g (group 0: a_n of
        // odd n, b_n of even n; group 1 the other way), so that a lane's role
        // and strides are fixed and the operands of a step are plain loads
        const bool ap = lk < 2;
        const int c0 = 2 * (lk & 1) + (ap ? 1 : 2), c1 = 2 * (lk & 1) + (ap ? 2 : 1);
        const int ub = ap ? 0 : nph;
        const R *u0 = U + 16 * rt + li + (size_t)(ub + c0 - 1) * ld;
        const R *u1 = U + 16 * rt + li + (size_t)(ub + c1 - 1) * ld;
        // pp, pt: c_n pi_n and c_n tau_n of this lane's angle at n = 1, rows
        // tstride apart; x takes pi_n for a_n and tau_n for b_n, y the other
        auto run = [&](const R *pp, const R *pt, int tstride) {
          const R *px = ap ? pp : pt, *py = ap ? pt : pp;
#pragma unroll 4
          for (int sk = 0; sk < nsk; sk++) {
            const int n0 = 4 * sk + c0, n1 = 4 * sk + c1;
            const R a0 = n0 <= nrow ? u0[(size_t)(4 * sk) * ld] : (R)0;
            const R a1 = n1 <= nrow ? u1[(size_t)(4 * sk) * ld] : (R)0;
            const bool t0 = ang_ok && n0 <= nterms, t1 = ang_ok && n1 <= nterms;
            const R tx0 = t0 ? px[(n0 - 1) * tstride] : (R)0;
            const R ty0 = t0 ? py[(n0 - 1) * tstride] : (R)0;
            const R tx1 = t1 ? px[(n1 - 1) * tstride] : (R)0;
            const R ty1 = t1 ? py[(n1 - 1) * tstride] : (R)0;
            x0 = M2::mfma(a0, tx0, x0);
            y0 = M2::mfma(a0, ty0, y0);
            x1 = M2::mfma(a1, tx1, x1);
            y1 = M2::mfma(a1, ty1, y1);
          }
        };
        if (tab_lds)
          run(tab + hix, tab + TABN + hix, nh2);
        else if (tab_tile)
          run(tab + li, tab + TABN + li, 16);
        else
          run(tpi + ang, ttau + ang, nquad);
#pragma unroll
        for (int h2 = 0; h2 < 2; h2++) {
          const int bw = 8 * rt + M2::out_bin(lk, h2);
          const PA wgt = (bw < nbin && L.nmb[bw] > 0) ? (PA)(nd[bw] * dd[bw]) : (PA)0.0;
          const int re = 2 * h2, im = 2 * h2 + 1;
          // +mu
          R s1r = x0[re] + x1[re], s1i = x0[im] + x1[im];
          R s2r = y1[re] + y0[re], s2i = y1[im] + y0[im];
          R a1 = s1r * s1r + s1i * s1i, a2 = s2r * s2r + s2i * s2i;
          p1 += (PA)((R)0.5 * (a2 + a1)) * wgt;
          p2 += (PA)((R)0.5 * (a2 - a1)) * wgt;
          p3 += (PA)(s2r * s1r + s2i * s1i) * wgt;
          // -mu
          s1r = x0[re] - x1[re];
          s1i = x0[im] - x1[im];
          s2r = y1[re] - y0[re];
          s2i = y1[im] - y0[im];
          a1 = s1r * s1r + s1i * s1i;
          a2 = s2r * s2r + s2i * s2i;
          m1 += (PA)((R)0.5 * (a2 + a1)) * wgt;
          m2 += (PA)((R)0.5 * (a2 - a1)) * wgt;
          m3 += (PA)(s2r * s1r + s2i * s1i) * wgt;
        }
      }
      // the four lane groups hold different bins of the same angle
      p1 += __shfl_xor(p1, 16, 64);
      p1 += __shfl_xor(p1, 32, 64);
      p2 += __shfl_xor(p2, 16, 64);
      p2 += __shfl_xor(p2, 32, 64);
      p3 += __shfl_xor(p3, 16, 64);
      p3 += __shfl_xor(p3, 32, 64);
      m1 += __shfl_xor(m1, 16, 64);
      m1 += __shfl_xor(m1, 32, 64);
      m2 += __shfl_xor(m2, 16, 64);
      m2 += __shfl_xor(m2, 32, 64);
      m3 += __shfl_xor(m3, 16, 64);
      m3 += __shfl_xor(m3, 32, 64);
      if (lk == 0 && ang_ok) {
        if (wpc == 1) {  // this wave alone owns the pair
          L.ps[0][ang] = p1;
          L.ps[1][ang] = p2;
          L.ps[2][ang] = p3;
          if (angm != ang) {
            L.ps[0][angm] = m1;
            L.ps[1][angm] = m2;
            L.ps[2][angm] = m3;
          }
        } else {
          R *pw = L.pw[wv];
          pw[li] = p1;
          pw[16 + li] = p2;
          pw[32 + li] = p3;
          pw[48 + li] = m1;
          pw[64 + li] = m2;
          pw[80 + li] = m3;
        }
      }
    }
    __syncthreads();
    // the shared tiles, waves in order (with the table staged by tile,
    // the one tile of this pass)
    if (wpc > 1 && t < (tab_tile ? 16 : 16 * nct)) {
      const int ctw = tab_tile ? pass : t >> 4, l16 = t & 15;
      const int hix = tab_tile ? 16 * pass + t : t;
      if (hix < nh2) {
        const int ang = j0 + hix, angm = nquad - 1 - ang;
        double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int q = 0; q < wpc; q++) {
          const R *pw = L.pw[tab_tile ? q : ctw * wpc + q];
#pragma unroll
          for (int k = 0; k < 6; k++) v[k] += pw[16 * k + l16];
        }
        L.ps[0][ang] = v[0];
        L.ps[1][ang] = v[1];
        L.ps[2][ang] = v[2];
        if (angm != ang) {
          L.ps[0][angm] = v[3];
          L.ps[1][angm] = v[4];
          L.ps[2][angm] = v[5];
        }
      }
    }
    }  // pass
    // efficiency sums, waves in order
    double sumqe = 0.0, sumqs = 0.0, sumqb = 0.0;
    for (int q = 0; q < MIE2_NW; q++) {
      sumqe = sumqe + L.red[q][0];
      sumqs = sumqs + L.red[q][1];
      sumqb = sumqb + L.red[q][2];
    }
    if (t == 0) {
      long long ns_h = 0;
      for (int q = 0; q < MIE2_NW; q++) {
        work_bins += L.redi[q][0];
        ns_h += L.redi[q][1];
      }
      work_series += ns_h;
      work_ang += (long long)nquad * ns_h;
      work_proj += (long long)nquad * (nlegen + 1);
      L.nlh = 0;
      L.nan_ = 0;
    }
    const double extinction = s.lhyd_abs ? sumqe : sumqe - sumqs;
    const double scatter = sumqs;
    const double albedo = scatter / extinction;
    if (!(extinction > 0.0) || !(scatter > 0.0) || !(sumqb > 0.0)) {
      if (t == 0) set_err(st, E_MIE_RESULT);
      return;
    }
    __syncthreads();  // phase sums complete
    // weighted phase sums per angle, normalised (mie_spheres.f90:272-280), in place
    const double tmp = wavelength * wavelength / (PI * scatter);
    for (int k = t; k < nquad; k += MIE2_WG) {
      const double wt = twt[k];
      L.ps[0][k] = tmp * L.ps[0][k] * wt;
      L.ps[1][k] = tmp * L.ps[1][k] * wt;
      L.ps[2][k] = tmp * L.ps[2][k] * wt;
    }
    __syncthreads();
    // --- C: Legendre projection (mie_spheres.f90:272-314), thread per
    // order, over the pairs of angles: P_l(-mu) = (-1)^l P_l(mu)
    double g[3][MIE2_NL];
#pragma unroll
    for (int j = 0; j < MIE2_NL; j++) {
      const int l = t + MIE2_WG * j;
      g[0][j] = g[1][j] = g[2][j] = 0.0;
      if (l <= nlegen) {
        double c1 = 0.0, c2 = 0.0, c3 = 0.0;
        const double sg = (l & 1) ? -1.0 : 1.0;
#pragma unroll 4
        for (int h = 0; h < nh2; h++) {
          const int k = j0 + h, km = nquad - 1 - k;
          const double pl = tpl[(size_t)k * (nlegen + 1) + l];
          double v1 = L.ps[0][k], v2 = L.ps[1][k], v3 = L.ps[2][k];
          if (km != k) {
            v1 += sg * L.ps[0][km];
            v2 += sg * L.ps[1][km];
            v3 += sg * L.ps[2][km];
          }
          c1 += v1 * pl;
          c2 += v2 * pl;
          c3 += v3 * pl;
        }
        const double hl = (double)((float)(2 * l + 1) / 2.0f);
        g[0][j] = hl * c1;
        g[1][j] = hl * c2;
        g[2][j] = hl * c3;
        // the reference keeps the orders with a coefficient above 1e-7
        // (mie_spheres.f90:305-313); from FP32 S1/S2 the projection carries a
        // noise of ~1e-7 of g_0 itself, so the mixed path keeps those above
        // 1e-6 (the omitted orders change the phase function by < 1e-6 of
        // its mean)
        if (g[0][j] > (F32 ? 1.0e-6 : sp(1.0e-7))) atomicMax(&L.nlh, l);
        if (isnan(g[0][j]) || isnan(c2) || isnan(c3)) L.nan_ = 1;
      }
    }
    if (t == 0 && s.fault == FAULT_NAN_COEF && col == s.fault_col && f == s.fault_ch) L.nan_ = 1;
    __syncthreads();
    if (L.nan_) {
      if (t == 0) set_err(st, E_NAN_COEF);
      return;
    }
    // accumulate over hydrometeors (scatProperties.f90:920-946)
    const int nl_h = L.nlh;
    if (nl_h > nleg_coef) nleg_coef = nl_h;
    kexttot = kexttot + extinction;
    backtot = backtot + sumqb;  // rt_back (scatProperties.f90:743-757, 922)
    salbedo = (kexttot == 0.0) ? 0.0 : salbedo + (albedo * extinction);
#pragma unroll
    for (int j = 0; j < MIE2_NL; j++) {
      const int l = t + MIE2_WG * j;
      if (l < nleg_coef && l <= nlegen) {
        L.lc[0][l] = L.lc[0][l] + (g[0][j] * albedo * extinction);
        L.lc[1][l] = L.lc[1][l] + (g[1][j] * albedo * extinction);
        L.lc[2][l] = L.lc[2][l] + (g[2][j] * albedo * extinction);
      }
    }
    if (kexttot < 0.0 || isnan(kexttot)) {
      if (t == 0) set_err(st, E_KEXT);
      return;
    }
    if (nleg_coef > MAXNLEG - 1) {
      if (t == 0) set_err(st, E_NLEG);
      return;
    }
    __syncthreads();  // the next hydrometeor rewrites LDS and the slot
  }
  // finalize_rt3_scatProperties and output
  salbedo = salbedo / kexttot;
  const double norm = salbedo * kexttot;
  double *out = w.coef + (size_t)itf * NCOEF * MAXNLEG;
#pragma unroll
  for (int j = 0; j < MIE2_NL; j++) {
    const int l = t + MIE2_WG * j;
    if (l <= nleg_coef && l < MAXNLEG) {
      out[l] = L.lc[0][l] / norm;
      out[MAXNLEG + l] = L.lc[1][l] / norm;
      out[2 * MAXNLEG + l] = L.lc[2][l] / norm;
    }
  }
  if (t == 0) {
    w.kext[itf] = kexttot;
    w.back[itf] = backtot;
    w.salb[itf] = salbedo;
    w.nleg[itf] = nleg_coef;
    atomicAdd(&w.counters[6], (unsigned long long)work_bins);
    atomicAdd(&w.counters[7], (unsigned long long)work_series);
    atomicAdd(&w.counters[12], (unsigned long long)work_ang);
    atomicAdd(&w.counters[13], (unsigned long long)work_proj);
  }
}

// Persistent grid of the coefficient-first Mie: workgroup b owns slot b of
// the U scratch and takes units from the work queue until it is drained.
// Refractive indices of liquid water and ice (scatProperties.f90:227-243)
// per (layer, channel) unit, for the phases present in it: one thread per
// unit, so that their long serial evaluation overlaps across the whole grid
// instead of holding a workgroup of the coefficient-first kernel.
__global__ void k_mie_index(Dims d, const pam_hydro_desc *hyd, const double *freqs, Work w) {
  const int n = d.nitems * d.nf;
  for (int itf = blockIdx.x * blockDim.x + threadIdx.x; itf < n; itf += gridDim.x * blockDim.x) {
    const int li_item = itf / d.nf, f = itf % d.nf;
    const int col = w.item_col[li_item], z = w.item_layer[li_item];
    bool nw = false, ni = false;
    for (int h = 0; h < d.nh; h++)
      if (w.dsd_flag[(size_t)li_item * d.nh + h] > 0) {
        if (hyd[h].liq_ice == 1)
          nw = true;
        else
          ni = true;
      }
    const double T = w.temp[(size_t)col * d.ML + z], freq = freqs[f];
    cd mw = cd{0.0, 0.0}, mi = cd{0.0, 0.0};
    if (nw) mw = csqrt_(eps_water_tkc(T - sp(273.15), freq));
    if (ni) mi = csqrt_(eps_ice(T, freq));
    double *o = w.mie_m + (size_t)itf * 4;
    o[0] = mw.re;
    o[1] = mw.im;
    o[2] = mi.re;
    o[3] = mi.im;
  }
}

void launch_mie_index(const Dims &d, const pam_hydro_desc *hyd, const double *freqs, Work &w,
                      hipStream_t st) {
  const size_t n = (size_t)d.nitems * d.nf;
  if (n == 0) return;
  hipLaunchKernelGGL(k_mie_index, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)),
                     dim3(256), 0, st, d, hyd, freqs, w);
}

template <typename R>
__global__ void __launch_bounds__(MIE2_WG) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 3))) k_mie2(Dims d, Settings s, const pam_hydro_desc *hyd,
                                                   const double *freqs, Work w) {
  __shared__ Mie2Lds<R> L;
  R *U = reinterpret_cast<R *>(w.mie_u + (size_t)blockIdx.x * w.mie_slot_doubles);
  const int n = d.nitems * d.nf;
  for (int itf = next_unit(&w.queue[0], &L.slot); itf < n; itf = next_unit(&w.queue[0], &L.slot))
    mie2_unit<R>(d, s, hyd, freqs, w, itf, U, L);
}

// One unit per workgroup, or (PERS) a persistent grid taking units from a
// work queue (separate instantiations).
template <bool PERS, bool SSRG>
__global__ void __launch_bounds__(MIE_WG) __attribute__((amdgpu_waves_per_eu(2))) k_mie(
    Dims d, Settings s, const pam_hydro_desc *hyd, Hydro4D h4, const double *freqs, Work w) {
  __shared__ MieLds L;
  const int n = d.nitems * d.nf;
  int itf = PERS ? next_unit(&w.queue[0], &L.slot) : (int)blockIdx.x;
  while (itf < n) {
    mie_unit<SSRG>(d, s, hyd, h4, freqs, w, itf, L);
    if (!PERS) break;
    itf = next_unit(&w.queue[0], &L.slot);
  }
}

int mie_slot_rows(int nbmax) { return 16 * ((nbmax + 7) / 8); }
size_t mie_slot_doubles(int nbmax) { return (size_t)mie_slot_rows(nbmax) * 2 * MAX_NTERMS; }

void launch_mie(const Dims &d, const Settings &s, const pam_hydro_desc *hyd, const Hydro4D &h4,
                const Profile &p, const double *freqs, Work &w, hipStream_t st) {
  (void)p;
  size_t n = (size_t)d.nitems * d.nf;
  if (n == 0) return;
  if (!s.ssrg && !s.mie_chunked && w.mie_u && w.mie_slots > 0) {
    const unsigned grid = (unsigned)std::min<size_t>(n, (size_t)w.mie_slots);
    hipLaunchKernelGGL(k_mie_index, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)),
                       dim3(256), 0, st, d, hyd, freqs, w);
    if (s.precision == PAM_PREC_MIXED && !s.mie_fp64 && w.atab_f)
      hipLaunchKernelGGL(k_mie2<float>, dim3(grid), dim3(MIE2_WG), 0, st, d, s, hyd, freqs, w);
    else
      hipLaunchKernelGGL(k_mie2<double>, dim3(grid), dim3(MIE2_WG), 0, st, d, s, hyd, freqs, w);
    return;
  }
  const dim3 grid((unsigned)persistent_grid(n, s.pers_mie, s.ncu));
  if (s.ssrg) {
    if (grid.x < n)
      hipLaunchKernelGGL((k_mie<true, true>), grid, dim3(MIE_WG), 0, st, d, s, hyd, h4, freqs,
                         w);
    else
      hipLaunchKernelGGL((k_mie<false, true>), grid, dim3(MIE_WG), 0, st, d, s, hyd, h4, freqs,
                         w);
  } else if (grid.x < n) {
    hipLaunchKernelGGL((k_mie<true, false>), grid, dim3(MIE_WG), 0, st, d, s, hyd, h4, freqs,
                         w);
  } else {
    hipLaunchKernelGGL((k_mie<false, false>), grid, dim3(MIE_WG), 0, st, d, s, hyd, h4, freqs,
                         w);
  }
}

}  // namespace pam
