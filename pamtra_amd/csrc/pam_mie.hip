// pam_mie.hip -- Mie single scattering of a layer's hydrometeors and the
// Legendre expansion of the layer phase function: calc_mie_spheres
// (mie_spheres.f90:14-345) with miecalc/miecross/mieangle
// (mie_scat_utilities.f90:8-160), accumulated over hydrometeors as
// calc_scatProperties does (scatProperties.f90:920-946) and normalised as in
// finalize_rt3_scatProperties (:993-1002).
//
// One 64-lane wavefront per (scattering layer, channel).  Lanes run the
// per-bin Mie recurrences in parallel (bins in chunks, coefficients staged in
// LDS), then own quadrature angles for the DSD-weighted phase-function sums,
// which stay in registers; the Legendre projection is a wavefront reduction
// per order.  Bin sums follow the reference's bin order.
#include "pam_internal.h"

namespace pam {

constexpr int MIE_WG = 64;
constexpr int AB_CAP = 1024;  // (a_n, b_n) pairs staged in LDS per chunk
constexpr int MAXK = MAX_NQUAD / MIE_WG;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

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

// Sihvola mixing with air applied to a refractive index (eps_mix.f90:1-48,
// called as in mie_spheres.f90:187-192); single-precision result
__device__ cd eps_mix_air(cd e1, double rho) {
  const double v = 0.85;
  double f = rho / 917.0;
  double r1 = e1.re;
  double b = r1 + 2.0 - 2.0 * v - f * (r1 - 1.0) * (1.0 + v);
  double a = (r1 + (2.0 - v) + f * (2.0 - v) * (r1 - 1.0));
  double emr = (sqrt(b * b + 4.0 * v * a) - b) / (2.0 * v);
  double emi = e1.im * (-(emr - 1.0) + f * (emr + 2.0 + v * (emr - 1.0))) /
               ((r1 + 2.0 + 2.0 * v * (emr - 1.0)) - f * (1.0 + v) * (r1 - 1.0));
  return cd{sp(emr), sp(emi)};
}

__device__ __forceinline__ int mie_nstop(double x) {
  return (int)(x + 4.0 * pow(x, sp(0.3334)) + 2.0);
}

__global__ void __launch_bounds__(MIE_WG) k_mie(Dims d, Settings s, const pam_hydro_desc *hyd,
                                                const double *freqs, Work w) {
  __shared__ double s_mu[MAX_NQUAD], s_wt[MAX_NQUAD];
  __shared__ cd s_a[AB_CAP], s_b[AB_CAP];
  __shared__ double s_q[3][MIE_WG];
  __shared__ int s_nm[MIE_WG];
  __shared__ double s_leg[NCOEF][MAXNLEG];
  __shared__ double s_lc[NCOEF][MAXNLEG + 1];

  const int lane = threadIdx.x;
  const int itf = blockIdx.x;
  const int li = itf / d.nf, f = itf % d.nf;
  const int col = w.item_col[li], z = w.item_layer[li];
  int32_t *st = w.itf_status + itf;
  const size_t oy = (size_t)col * d.ML + z;
  const double T = w.temp[oy];
  const double freq = freqs[f];
  const double wavelength = C_LIGHT / (freq * 1e9);

  for (int k = lane; k < NCOEF * (MAXNLEG + 1); k += MIE_WG) (&s_lc[0][0])[k] = 0.0;
  double kexttot = 0.0, salbedo = 0.0;
  int nleg_coef = 0;
  long long work_bins = 0, work_series = 0;

  for (int hh = 0; hh < d.nh; hh++) {
    const int nbin = w.dsd_flag[(size_t)li * d.nh + hh];
    if (nbin <= 0) continue;
    const pam_hydro_desc &hd = hyd[hh];
    const size_t ob = ((size_t)li * d.nh + hh) * d.nbmax;
    const double *dia = w.dsd_d + ob, *dd = w.dsd_dd + ob, *nd = w.dsd_nd + ob,
                 *rho = w.dsd_rho + ob;
    // refractive index (scatProperties.f90:227-243)
    cd m;
    if (hd.liq_ice == 1)
      m = csqrt_(eps_water_tkc(T - sp(273.15), freq));
    else
      m = csqrt_(eps_ice(T, freq));
    const cd m_ice = cd{m.re, -m.im};
    const bool is_ice = (hd.liq_ice == -1);
    // series length from the largest bin (mie_spheres.f90:139-158)
    const double xmax = PI * dia[nbin - 1] / wavelength;
    const int nterms = mie_nstop(xmax);
    int nlegen = 2 * nterms;
    if (nlegen > MAXNLEG - 1) nlegen = MAXNLEG - 1;
    const int nquad = (nlegen + 2 * nterms + 2) / 2;
    if (nterms > MAX_NTERMS || nquad > MAX_NQUAD) {
      if (lane == 0) set_err(st, E_MIE_TERMS);
      return;
    }
    // Gauss-Legendre nodes (scat_utilities.f90:1-44), ascending
    for (int j = lane + 1; j <= (nquad + 1) / 2; j += MIE_WG) {
      double x = cos(PI * (j - 0.25) / (nquad + 0.5));
      double pl, pl1, pl2, dpl, xp;
      int it = 0;
      do {
        pl1 = 1.0;
        pl = x;
        for (int l = 2; l <= nquad; l++) {
          pl2 = pl1;
          pl1 = pl;
          pl = ((double)(2 * l - 1) * x * pl1 - (double)(l - 1) * pl2) / (double)l;
        }
        dpl = (double)nquad * (x * pl - pl1) / (x * x - 1.0);
        xp = x;
        x = xp - pl / dpl;
        it++;
      } while (fabs(x - xp) > 3.0e-13 && it < 10);
      double wt = 2.0 / ((1.0 - x * x) * dpl * dpl);
      s_mu[j - 1] = -x;
      s_wt[j - 1] = wt;
      s_mu[nquad - j] = x;
      s_wt[nquad - j] = wt;
    }
    __syncthreads();
    double mu_k[MAXK];
#pragma unroll
    for (int k = 0; k < MAXK; k++) {
      int i = lane + MIE_WG * k;
      mu_k[k] = (i < nquad) ? s_mu[i] : 0.0;
    }
    double sp1[MAXK], sp2[MAXK], sp3[MAXK];
#pragma unroll
    for (int k = 0; k < MAXK; k++) sp1[k] = sp2[k] = sp3[k] = 0.0;
    double sumqe = 0.0, sumqs = 0.0, sumqb = 0.0;

    const int chunk = min(MIE_WG, AB_CAP / nterms);
    for (int b0 = 0; b0 < nbin; b0 += chunk) {
      const int nb = min(chunk, nbin - b0);
      // --- lanes over bins: a_n, b_n by miecalc, efficiencies by miecross
      if (lane < nb) {
        const int ir = b0 + lane;
        double q3[3] = {0.0, 0.0, 0.0};
        int nmie = 0;
        if (nd[ir] > 0.0) {
          // soft sphere: index of the ice-air mixture at the bin's density
          // (mie_spheres.f90:187-192), decided per bin
          cd msp = (is_ice && rho[ir] != 917.0) ? eps_mix_air(m_ice, rho[ir]) : m_ice;
          double x = PI * dia[ir] / wavelength;
          nmie = mie_nstop(x);
          if (nmie > nterms) {
            set_err(st, E_MIE_TERMS);
            nmie = nterms;
          }
          cd mc = conj(msp);
          cd y = mc * x;
          cd *A = s_a + lane * nterms, *B = s_b + lane * nterms;
          // downward log-derivative recursion, D_n kept in the a-slot
          int nn = nmie + 15;
          cd dn = cd{0.0, 0.0};
          for (int n = nn; n >= 2; n--) {
            cd ny = cdiv(cd{(double)n, 0.0}, y);
            dn = ny - cdiv(cd{1.0, 0.0}, dn + ny);
            if (n - 1 <= nmie) A[n - 2] = dn;
          }
          double psim = cos(x), psin = sin(x), chim = -sin(x), chin = cos(x);
          double s1 = 0.0, s2 = 0.0;
          cd s3 = cd{0.0, 0.0};
          for (int n = 1; n <= nmie; n++) {
            double tmp = psin;
            psin = (double)(2 * n - 1) / x * psin - psim;
            psim = tmp;
            tmp = chin;
            chin = (double)(2 * n - 1) / x * chin - chim;
            chim = tmp;
            cd xin = cd{sp(psin), sp(-chin)};
            cd xim = cd{sp(psim), sp(-chim)};
            cd Dn = A[n - 1];
            cd ct = cdiv(Dn, mc) + (double)n / x;
            cd an = cdiv(ct * psin - psim, ct * xin - xim);
            ct = mc * Dn + (double)n / x;
            cd bn = cdiv(ct * psin - psim, ct * xin - xim);
            A[n - 1] = an;
            B[n - 1] = bn;
            double wn = (double)(2 * n + 1);
            s1 = s1 + wn * (an.re + bn.re);
            s2 = s2 + wn * ((an * conj(an)).re + (bn * conj(bn)).re);
            double sg = (n & 1) ? -1.0 : 1.0;
            s3 = s3 + (wn * sg) * (an - bn);
          }
          double r = dia[ir] / 2.0;
          double ar = (r * r) * PI;
          double qe = 2.0 / (x * x) * s1, qs = 2.0 / (x * x) * s2;
          double qb = ((1.0 / (x * x) * s3) * conj(s3)).re;
          q3[0] = qe * ar * nd[ir] * dd[ir];
          q3[1] = qs * ar * nd[ir] * dd[ir];
          q3[2] = qb * ar * nd[ir] * dd[ir];
        }
        s_q[0][lane] = q3[0];
        s_q[1][lane] = q3[1];
        s_q[2][lane] = q3[2];
        s_nm[lane] = nmie;
      }
      __syncthreads();
      // --- lanes over angles: DSD-weighted phase-function sums
      for (int cb = 0; cb < nb; cb++) {
        const int nmie = s_nm[cb];
        if (nmie <= 0) continue;
        const int ir = b0 + cb;
        const double wgt = nd[ir] * dd[ir];
        const cd *A = s_a + cb * nterms, *B = s_b + cb * nterms;
#pragma unroll
        for (int k = 0; k < MAXK; k++) {
          if (lane + MIE_WG * k >= nquad) break;
          const double mu = mu_k[k];
          cd S1 = cd{0.0, 0.0}, S2 = cd{0.0, 0.0};
          double pin = 1.0, pim = 0.0;
          for (int n = 1; n <= nmie; n++) {
            double taun = (double)n * mu * pin - (double)(n + 1) * pim;
            double c = (double)((float)(2 * n + 1) / (float)(n * (n + 1)));
            cd an = A[n - 1], bn = B[n - 1];
            S1 = S1 + c * (an * pin + bn * taun);
            S2 = S2 + c * (bn * pin + an * taun);
            double tmp = pin;
            pin = ((double)(2 * n + 1) * mu * pin - (double)(n + 1) * pim) / (double)n;
            pim = tmp;
          }
          double a1 = cabs_(S1), a2 = cabs_(S2);
          double p1 = 0.5 * (a2 * a2 + a1 * a1);
          double p2 = 0.5 * (a2 * a2 - a1 * a1);
          double p3 = (conj(S2) * S1).re;
          sp1[k] = sp1[k] + p1 * wgt;
          sp2[k] = sp2[k] + p2 * wgt;
          sp3[k] = sp3[k] + p3 * wgt;
        }
        if (lane == 0) {
          sumqe = sumqe + s_q[0][cb];
          sumqs = sumqs + s_q[1][cb];
          sumqb = sumqb + s_q[2][cb];
          work_bins++;
          work_series += nmie;
        }
      }
      __syncthreads();
    }
    // broadcast the ordered bin sums of lane 0
    sumqe = __shfl(sumqe, 0, 64);
    sumqs = __shfl(sumqs, 0, 64);
    sumqb = __shfl(sumqb, 0, 64);
    const double extinction = s.lhyd_abs ? sumqe : sumqe - sumqs;
    const double scatter = sumqs;
    const double albedo = scatter / extinction;
    if (!(extinction > 0.0) || !(scatter > 0.0) || !(sumqb > 0.0)) {
      if (lane == 0) set_err(st, E_MIE_RESULT);
      return;
    }
    // Legendre projection (mie_spheres.f90:272-314)
    const double tmp = wavelength * wavelength / (PI * scatter);
    double pl[MAXK], pl1[MAXK], pl2[MAXK];
#pragma unroll
    for (int k = 0; k < MAXK; k++) {
      int i = lane + MIE_WG * k;
      double wt = (i < nquad) ? s_wt[i] : 0.0;
      sp1[k] = tmp * sp1[k] * wt;
      sp2[k] = tmp * sp2[k] * wt;
      sp3[k] = tmp * sp3[k] * wt;
      pl[k] = 1.0;
      pl1[k] = 1.0;
      pl2[k] = 0.0;
    }
    int nl_h = 0;
    for (int l = 0; l <= nlegen; l++) {
      double c1 = 0.0, c2 = 0.0, c3 = 0.0;
#pragma unroll
      for (int k = 0; k < MAXK; k++) {
        if (l > 0)
          pl[k] = (double)(2 * l - 1) * mu_k[k] * pl1[k] / (double)l -
                  (double)(l - 1) * pl2[k] / (double)l;
        c1 += sp1[k] * pl[k];
        c2 += sp2[k] * pl[k];
        c3 += sp3[k] * pl[k];
        pl2[k] = pl1[k];
        pl1[k] = pl[k];
      }
      c1 = wave_sum(c1);
      c2 = wave_sum(c2);
      c3 = wave_sum(c3);
      const double hl = (double)((float)(2 * l + 1) / 2.0f);
      const double g1 = hl * c1;
      if (lane == 0) {
        s_leg[0][l] = g1;
        s_leg[1][l] = hl * c2;
        s_leg[2][l] = hl * c3;
      }
      if (g1 > sp(1.0e-7)) nl_h = l;
      if (isnan(g1) || isnan(c2) || isnan(c3)) {
        if (lane == 0) set_err(st, E_NAN_COEF);
        return;
      }
    }
    for (int l = nlegen + 1 + lane; l < MAXNLEG; l += MIE_WG) {
      s_leg[0][l] = 0.0;
      s_leg[1][l] = 0.0;
      s_leg[2][l] = 0.0;
    }
    __syncthreads();
    // accumulate over hydrometeors (scatProperties.f90:920-946)
    if (nl_h > nleg_coef) nleg_coef = nl_h;
    kexttot = kexttot + extinction;
    salbedo = (kexttot == 0.0) ? 0.0 : salbedo + (albedo * extinction);
    for (int jj = lane; jj < nleg_coef; jj += MIE_WG) {
      s_lc[0][jj] = s_lc[0][jj] + (s_leg[0][jj] * albedo * extinction);
      s_lc[1][jj] = s_lc[1][jj] + (s_leg[1][jj] * albedo * extinction);
      s_lc[2][jj] = s_lc[2][jj] + (s_leg[2][jj] * albedo * extinction);
    }
    __syncthreads();
    if (kexttot < 0.0 || isnan(kexttot)) {
      if (lane == 0) set_err(st, E_KEXT);
      return;
    }
    if (nleg_coef > MAXNLEG - 1) {
      if (lane == 0) set_err(st, E_NLEG);
      return;
    }
  }
  // finalize_rt3_scatProperties and output
  salbedo = salbedo / kexttot;
  const double norm = salbedo * kexttot;
  double *out = w.coef + (size_t)itf * NCOEF * MAXNLEG;
  for (int l = lane; l <= nleg_coef && l < MAXNLEG; l += MIE_WG) {
    out[l] = s_lc[0][l] / norm;
    out[MAXNLEG + l] = s_lc[1][l] / norm;
    out[2 * MAXNLEG + l] = s_lc[2][l] / norm;
  }
  if (lane == 0) {
    w.kext[itf] = kexttot;
    w.salb[itf] = salbedo;
    w.nleg[itf] = nleg_coef;
    atomicAdd(&w.counters[6], (unsigned long long)work_bins);
    atomicAdd(&w.counters[7], (unsigned long long)work_series);
  }
}

void launch_mie(const Dims &d, const Settings &s, const pam_hydro_desc *hyd, const Hydro4D &h4,
                const Profile &p, const double *freqs, Work &w, hipStream_t st) {
  (void)h4;
  (void)p;
  size_t n = (size_t)d.nitems * d.nf;
  if (n == 0) return;
  hipLaunchKernelGGL(k_mie, dim3((unsigned)n), dim3(MIE_WG), 0, st, d, s, hyd, freqs, w);
}

}  // namespace pam
