// pam_dsd.hip -- drop-size distributions on the device, one thread per
// (scattering layer, hydrometeor): mass-size law (make_mass_size.f90:30-148),
// distribution parameters (make_dist_param.f90:2-631), size grid and bin
// numbers (make_dist.f90:2-290) and soft-sphere density
// (make_soft_spheroid.f90:78-152).  The adaptive 500-point scan is streamed:
// the merged linear/log grid is generated on the fly and only the two
// crossing points are kept, so a thread needs no per-point storage.
#include "pam_internal.h"

namespace pam {

struct HP {
  double as_ratio, rho_ms, a_ms, b_ms, p_1, p_2, p_3, p_4, d_1, d_2;
  int liq_ice, moment_in, nbin, dist;
};

struct DP {
  double n_0, lambda, mu, gam, n_t, sig, d_ln, d_mono;
};

__device__ __forceinline__ double h4val(const double *a4, double v, size_t i) {
  return a4 ? a4[i] : v;
}

// make_mass_size (make_mass_size.f90:30-148); returns an error code
__device__ int32_t mass_size(HP &h) {
  if (h.b_ms > 3.5 || h.a_ms == 0.0 || h.rho_ms == 0.0 || h.as_ratio == 0.0 ||
      (h.liq_ice == -1 && h.rho_ms > 917.0))
    return E_MASS_SIZE;
  if (h.liq_ice != 1) {
    if (h.liq_ice == -1 && h.a_ms < 0.0 && h.b_ms < 0.0 && h.rho_ms < 0.0) return E_MASS_SIZE;
    if ((h.a_ms < 0.0 && h.b_ms >= 0.0) || (h.a_ms > 0.0 && h.b_ms < 0.0)) return E_MASS_SIZE;
  }
  if (h.liq_ice == 1) {
    h.b_ms = 3.0;
    h.a_ms = PI / 6.0 * 1000.0;
    return 0;
  }
  if (h.a_ms > 0.0 && h.b_ms >= 0 && h.liq_ice == -1) return 0;
  if (h.liq_ice == -1 && (h.as_ratio < 0.0 || h.as_ratio == 1.0)) {
    h.b_ms = 3.0;
    h.a_ms = PI / 6.0 * h.rho_ms;
    return 0;
  }
  if (h.liq_ice == -1 && h.as_ratio > 1.0) {
    h.b_ms = 3.0;
    h.a_ms = PI / 6.0 * h.rho_ms * pow(1.0 / h.as_ratio, 2.0);
    return 0;
  }
  if (h.liq_ice == -1 && h.as_ratio > 0.0 && h.as_ratio < 1.0) {
    h.b_ms = 3.0;
    h.a_ms = PI / 6.0 * h.rho_ms * h.as_ratio;
    return 0;
  }
  return E_MASS_SIZE;
}

// make_dist_params (make_dist_param.f90) for mono, exp, exp_field_t,
// exp_ryan, mgamma and logn
__device__ int32_t dist_params(const HP &h, double q_h, double n_tot, double r_eff,
                               double layer_t, DP &o) {
  const double NaN = __builtin_nan("");
  o.n_0 = -99.;
  o.lambda = -99.;
  o.gam = NaN;
  o.mu = NaN;
  o.n_t = -99.;
  o.sig = -99.;
  o.d_ln = NaN;
  o.d_mono = -99.;
  int mi = h.moment_in;
  if ((mi == 3 && q_h <= 0.) || (mi == 1 && n_tot <= 0.) || (mi == 2 && r_eff <= 0.) ||
      (mi == 13 && (q_h <= 0. || n_tot <= 0.)) || (mi == 23 && (q_h <= 0. || r_eff <= 0.)) ||
      (mi == 12 && (n_tot <= 0. || r_eff <= 0.)))
    return E_DIST_PARAMS;
  const double a = h.a_ms, b = h.b_ms, p1 = h.p_1, p2 = h.p_2, p3 = h.p_3, p4 = h.p_4;
  switch (h.dist) {
    case PAM_DIST_MONO: {
      o.lambda = 0.0;
      o.mu = 0.0;
      o.gam = 0.0;
      if (h.d_1 != -99. && p1 == -99.) {
        o.d_mono = h.d_1;
        if (mi == 3) o.n_0 = q_h / (DELTA_D_MONO * a * pow(h.d_1, b));
        if (mi == 1) o.n_0 = n_tot / DELTA_D_MONO;
      }
      if (p1 != -99. && h.d_1 == -99.) {
        o.n_0 = p1 / DELTA_D_MONO;
        if (mi == 3) o.d_mono = pow(q_h / (p1 * a), 1.0 / b);
        if (mi == 2) o.d_mono = r_eff * 2.0;
      }
      if (mi == 23) {
        o.d_mono = r_eff * 2.0;
        o.n_0 = q_h / (DELTA_D_MONO * a * pow(o.d_mono, b));
      }
      return (o.n_0 <= 0.0 || o.d_mono <= 0.0) ? E_DIST_PARAMS : 0;
    }
    case PAM_DIST_EXP:
    case PAM_DIST_EXP_FIELD_T:
    case PAM_DIST_EXP_RYAN: {
      o.gam = 1.0;
      o.mu = 0.0;
      double g = tgamma(b + 1.0);
      if (h.dist == PAM_DIST_EXP) {
        if (p1 != -99. && p2 != -99. && mi == 0) {
          o.lambda = p1;
          o.n_0 = p2;
        }
        if (p1 != -99. && p2 != -99. && mi == 3) {
          o.lambda = pow(p1 * a * g / q_h, 1.0 / (b + 1 - p2));
          o.n_0 = p1 * pow(o.lambda, p2);
        }
        if (p1 != -99. && p2 == -99. && p3 == -99.) {
          if (mi == 3) o.lambda = pow(p1 * a * g / q_h, 1.0 / b);
          if (mi == 2) o.lambda = 3.0 / r_eff;
          o.n_0 = p1 * o.lambda;
        }
        if (p1 == -99. && p2 != -99. && p3 == -99.) {
          o.lambda = 3.0 / p2;
          if (mi == 3) o.n_0 = (q_h * pow(o.lambda, b + 1.0)) / (a * g);
          if (mi == 1) o.n_0 = n_tot * o.lambda;
        }
        if (p1 == -99. && p2 == -99. && p3 != -99.) {
          o.n_0 = p3;
          if (mi == 3) o.lambda = pow(o.n_0 * a * g / q_h, 1.0 / (b + 1.0));
          if (mi == 1) o.lambda = o.n_0 / n_tot;
          if (mi == 2) o.lambda = 3.0 / r_eff;
        }
        if (p1 == -99. && p2 == -99.) {
          if (mi == 13) {
            o.lambda = pow(n_tot * a * g / q_h, 1.0 / b);
            o.n_0 = n_tot * o.lambda;
          }
          if (mi == 23) {
            o.lambda = 3.0 / r_eff;
            o.n_0 = (q_h * pow(o.lambda, b + 1)) / (a * g);
          }
          if (mi == 12) {
            o.lambda = 3.0 / r_eff;
            o.n_0 = n_tot * o.lambda;
          }
        }
      }
      if (h.dist == PAM_DIST_EXP_RYAN) {
        o.lambda = 1220.0 * pow(10.0, -0.0245 * (273.15 - layer_t));
        if (mi == 3) o.n_0 = (q_h * pow(o.lambda, b + 1)) / (a * g);
        if (mi == 1) o.n_0 = n_tot * o.lambda;
      }
      if (h.dist == PAM_DIST_EXP_FIELD_T) {
        o.n_0 = 7.628e6 * exp(0.107 * (273.15 - layer_t));
        if (mi == 3) o.lambda = pow(a * o.n_0 * g / q_h, 1.0 / (b + 1.0));
        if (mi == 2) o.lambda = 3.0 / r_eff;
      }
      return (o.n_0 <= 0.0 || o.lambda <= 0.0) ? E_DIST_PARAMS : 0;
    }
    case PAM_DIST_LOGN: {
      if (p1 != -99. && p3 != -99.) {
        o.n_t = p1;
        o.sig = p3;
        if (mi == 3) o.d_ln = (1.0 / b) * (log(q_h / (a * o.n_t)) - (0.5 * pow(b * o.sig, 2.0)));
        if (mi == 2) o.d_ln = log(2.0 * r_eff) - 2.5 * pow(o.sig, 2.0);
      }
      if (p1 != -99. && p2 != -99. && p3 == -99. && mi == 3) {
        o.n_t = p1;
        double w1 = q_h / (a * o.n_t * pow(2.0 * p2, 2.0));
        double w2 = 2.0 / (b * (b - 5.0));
        o.sig = sqrt(log(w1) * w2);
        o.d_ln = log(2.0 * p2) - 2.5 * pow(o.sig, 2.0);
      }
      if (p1 == -99. && p2 != -99. && p3 != -99.) {
        o.sig = p3;
        o.d_ln = log(2.0 * p2) - 2.5 * pow(o.sig, 2.0);
        if (mi == 3) o.n_t = q_h / a * exp(-1.0 * ((b * o.d_ln) + (0.5 * pow(b * o.sig, 2.0))));
        if (mi == 1) o.n_t = n_tot;
      }
      if (p1 == -99. && p2 == -99. && p3 != -99.) {
        o.sig = p3;
        if (mi == 13) {
          o.n_t = n_tot;
          o.d_ln = (1.0 / b) * (log(q_h / (a * o.n_t)) - (0.5 * pow(b * o.sig, 2.0)));
        }
        if (mi == 23) {
          o.d_ln = log(2.0 * r_eff) - 2.5 * pow(o.sig, 2.0);
          o.n_t = q_h / a * exp(-1.0 * ((b * o.d_ln) + (0.5 * pow(b * o.sig, 2.0))));
        }
        if (mi == 12) {
          o.n_t = n_tot;
          o.d_ln = log(2.0 * r_eff) - 2.5 * pow(o.sig, 2.0);
        }
      }
      return (o.n_t <= 0.0 || isnan(o.d_ln) || o.sig <= 0.0) ? E_DIST_PARAMS : 0;
    }
    case PAM_DIST_MGAMMA: {
      if (p3 == -99. || p4 == -99.) return E_DIST_PARAMS;
      double mu = p3, gam = p4;
      o.mu = mu;
      o.gam = gam;
      if (p1 != -99. && p2 == -99.) {
        if (mi == 3) {
          double w2 = tgamma((mu + b + 1.0) / gam), w3 = tgamma((mu + 1.0) / gam);
          o.lambda = pow(a / q_h * p1 * w2 / w3, gam / b);
          o.n_0 = gam * p1 / w3 * pow(o.lambda, (mu + 1.0) / gam);
        }
        if (mi == 2) {
          double w1 = tgamma((mu + 4.0) / gam), w2 = tgamma((mu + 3.0) / gam) * r_eff * 2.0;
          o.lambda = pow(w1 / w2, gam);
          w1 = (mu + 1.0) / gam;
          o.n_0 = p1 * gam * pow(o.lambda, w1) / tgamma(w1);
        }
      }
      if (p1 == -99. && p2 != -99.) {
        double w1 = tgamma((mu + 4.0) / gam), w2 = tgamma((mu + 3.0) / gam) * p2 * 2.0;
        if (mi == 3) {
          o.lambda = pow(w1 / w2, gam);
          double w1b = q_h * gam / a, w2b = (mu + b + 1.0) / gam;
          o.n_0 = w1b * pow(o.lambda, w2b) / tgamma(w2b);
        }
        if (mi == 1) {
          o.lambda = pow(w1 / w2, gam);
          double w1b = (mu + 1.0) / gam;
          o.n_0 = n_tot * gam * pow(o.lambda, w1b) / tgamma(w1b);
        }
      }
      if (p1 == -99. && p2 == -99.) {
        if (mi == 13) {
          double w2 = tgamma((mu + b + 1.0) / gam), w3 = tgamma((mu + 1.0) / gam);
          o.lambda = pow(a / q_h * n_tot * w2 / w3, gam / b);
          o.n_0 = gam * n_tot / w3 * pow(o.lambda, (mu + 1.0) / gam);
        }
        if (mi == 23) {
          double w1 = tgamma((mu + 4.0) / gam), w2 = tgamma((mu + 3.0) / gam) * r_eff * 2.0;
          o.lambda = pow(w1 / w2, gam);
          double w1b = q_h * gam / a, w2b = (mu + b + 1.0) / gam;
          o.n_0 = w1b * pow(o.lambda, w2b) / tgamma(w2b);
        }
        if (mi == 12) {
          double w1 = tgamma((mu + 4.0) / gam), w2 = tgamma((mu + 3.0) / gam) * r_eff * 2.0;
          o.lambda = pow(w1 / w2, gam);
          double w1b = (mu + 1.0) / gam;
          o.n_0 = n_tot * gam * pow(o.lambda, w1b) / tgamma(w1b);
        }
      }
      return (o.lambda <= 0.0 || o.n_0 <= 0.0) ? E_DIST_PARAMS : 0;
    }
    default:
      return E_DIST_PARAMS;
  }
}

__device__ __forceinline__ double dist_value(int dist, const DP &p, double d) {
  if (dist == PAM_DIST_LOGN)
    return p.n_t / (d * p.sig * sqrt(2.0 * PI)) *
           exp(-(pow(log(d) - p.d_ln, 2.0)) / (2.0 * (p.sig * p.sig)));
  return p.n_0 * pow(d, p.mu) * exp(-p.lambda * pow(d, p.gam));
}

// Generator of the merged linear/log bin-edge grid of make_dist.f90:120-155:
// nw+1 ascending edges on [d1, d2], the log list first on ties.
struct Grid {
  double d1, w_lin, w_log, logd1;
  int nlin, nlog, il, ig;
  __device__ void init(double a, double b, int nw) {
    nlin = (int)(nw / 2.0);
    nlog = (int)(nw / 2.0);
    if (nlin + nlog == nw) nlin = nlin - 1;
    d1 = a;
    w_lin = (b - a) / (nlin + 1);
    logd1 = log(a);
    w_log = (log(b) - logd1) / (nlog + 1);
    il = 0;
    ig = 0;
  }
  __device__ double next() {
    const double big = sp(1.e30);
    double ml = (il <= nlin) ? d1 + w_lin * (il + 1) : big;
    double mg = (ig <= nlog) ? exp(w_log * ig + logd1) : big;
    if (mg <= ml) {
      ig++;
      return mg;
    }
    il++;
    return ml;
  }
};

__global__ void __launch_bounds__(128) k_dsd(Dims d, Settings s, Profile p,
                                             const pam_hydro_desc *hyd, Hydro4D h4, Work w) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= d.nitems * d.nh) return;
  int li = idx / d.nh, hh = idx % d.nh;
  int col = w.item_col[li], z = w.item_layer[li];
  size_t oy = (size_t)col * d.ML + z;
  uint32_t mask = w.hmask[oy];
  int32_t *flag = w.dsd_flag + (size_t)li * d.nh + hh;
  *flag = 0;
  if (!(mask & (1u << hh))) return;
  int32_t *cst = w.col_status + col;
  const pam_hydro_desc &hd = hyd[hh];
  size_t i4 = oy * d.nh + hh;
  // as_ratio of the layer for the self-similar Rayleigh-Gans size along the
  // beam (scatProperties.f90:276, 286-287: negative or NaN reads as 0)
  const double ar4 = h4val(h4.as_ratio, hd.as_ratio, i4);
  w.dsd_ar[(size_t)li * d.nh + hh] = (ar4 < 0.0 || isnan(ar4)) ? 0.0 : ar4;
  if (h4.fs_nbin > 0) {
    // full spectrum given (hydrometeor_extinction.f90:120-150): the bins
    // replace the size distribution; the scattering diameter and density
    // are d_ds and rho_ds, delta_d the edge differences, ndens n / delta_d
    const int nb = h4.fs_nbin;
    const size_t b5 = i4 * (size_t)nb, b6 = i4 * (size_t)(nb + 1);
    const size_t o5 = ((size_t)li * d.nh + hh) * d.nbmax;
    double *fd = w.dsd_d + o5, *fdd = w.dsd_dd + o5, *fnd = w.dsd_nd + o5, *frho = w.dsd_rho + o5;
    const bool ssrg = hd.scat_name == PAM_SCAT_SSRG_RT3;
    for (int i = 0; i < nb; i++) {
      fd[i] = h4.fs_d[b5 + i];
      fdd[i] = h4.fs_d_bound[b6 + i + 1] - h4.fs_d_bound[b6 + i];
      fnd[i] = h4.fs_n[b5 + i] / fdd[i];
      frho[i] = ssrg ? h4.fs_mass[b5 + i] : h4.fs_rho[b5 + i];
    }
    for (int i = nb; i < d.nbmax; i++) {
      fd[i] = 1.0;
      fdd[i] = 1.0;
      fnd[i] = 0.0;
      frho[i] = 1000.0;
    }
    *flag = (hd.scat_name != PAM_SCAT_DISABLED) ? nb : 0;
    return;
  }
  HP h;
  h.as_ratio = h4val(h4.as_ratio, hd.as_ratio, i4);
  h.rho_ms = h4val(h4.rho_ms, hd.rho_ms, i4);
  h.a_ms = h4val(h4.a_ms, hd.a_ms, i4);
  h.b_ms = h4val(h4.b_ms, hd.b_ms, i4);
  h.p_1 = h4val(h4.p_1, hd.p_1, i4);
  h.p_2 = h4val(h4.p_2, hd.p_2, i4);
  h.p_3 = h4val(h4.p_3, hd.p_3, i4);
  h.p_4 = h4val(h4.p_4, hd.p_4, i4);
  h.d_1 = h4val(h4.d_1, hd.d_1, i4);
  h.d_2 = h4val(h4.d_2, hd.d_2, i4);
  h.liq_ice = hd.liq_ice;
  h.moment_in = hd.moment_in;
  h.dist = hd.dist_name;
  h.nbin = h4.nbin ? h4.nbin[i4] : hd.nbin;
  if (h.dist == PAM_DIST_MONO) h.nbin = 2;
  if (h.nbin > d.nbmax || h.nbin < 1) {
    set_err(cst, E_DIST_PARAMS);
    return;
  }
  // absolute moments (q2abs, conversions.f90:5-24)
  size_t b4 = oy * d.nh;
  double qsum = 0.0;
  bool qnan = false;
  for (int k = 0; k < d.nh; k++) {
    double q = p.hydro_q ? p.hydro_q[b4 + k] : __builtin_nan("");
    if (isnan(q)) qnan = true;
    qsum += q;
  }
  double T = w.temp[oy], P = w.press[oy], qv = w.q_hum[oy];
  double qall = (s.incl_hydro_rho && !qnan) ? qsum : 0.0;
  double den = R_D * (1.0 + (R_V / R_D - 1.0) * qv - qall) * T;
  double qin = p.hydro_q ? p.hydro_q[b4 + hh] : __builtin_nan("");
  double nin = p.hydro_n ? p.hydro_n[b4 + hh] : __builtin_nan("");
  double q_h = qin * P / den, n_tot = nin * P / den;
  double r_eff = p.hydro_reff ? p.hydro_reff[b4 + hh] : __builtin_nan("");

  int32_t e = mass_size(h);
  if (e) {
    set_err(cst, e);
    return;
  }
  DP dp;
  e = dist_params(h, q_h, n_tot, r_eff, T, dp);
  if (e) {
    set_err(cst, e);
    return;
  }
  const int nbin = h.nbin;
  const size_t ob = ((size_t)li * d.nh + hh) * d.nbmax;
  double *od = w.dsd_d + ob, *odd = w.dsd_dd + ob, *ond = w.dsd_nd + ob, *orho = w.dsd_rho + ob;
  const double thres = 0.1, min_d = sp(1.0e-7), max_d = sp(2.0e-2);
  double d1w, d2w;
  // pass A: edges, f at edges (kept in ond as scratch), centres, widths, mass
  double f_last = 0.0, mass_tot = 0.0;
  if (h.dist == PAM_DIST_MONO) {
    double bnd[3] = {dp.d_mono - DELTA_D_MONO * 0.5, dp.d_mono, dp.d_mono + DELTA_D_MONO * 0.5};
    double fv[3];
    for (int i = 0; i < 3; i++) fv[i] = dist_value(h.dist, dp, bnd[i]);
    for (int i = 0; i < 2; i++) {
      od[i] = (bnd[i] + bnd[i + 1]) * 0.5;
      odd[i] = bnd[i + 1] - bnd[i];
      ond[i] = fv[i];
      mass_tot = mass_tot + h.a_ms * (fv[i] + fv[i + 1]) / 2.0 * odd[i] * pow(od[i], h.b_ms);
    }
    f_last = fv[2];
  } else {
    if (s.adaptive) {
      // streamed adaptive scan (make_dist.f90:214-259)
      Grid g;
      g.init(min_d, max_d, 500);
      double bprev = g.next(), fprev = dist_value(h.dist, dp, bprev);
      double d2cross = -1.0, b1 = 0.0;
      bool have_b1 = false, done1 = false;
      double d1n = -1.0, bcur = bprev, fcur = fprev, bnw1 = bprev;
      for (int i = 1; i <= 500; i++) {
        bcur = g.next();
        fcur = dist_value(h.dist, dp, bcur);
        if (i == 1) b1 = bcur, have_b1 = true;
        if (fprev >= thres && fcur < thres) d2cross = bcur;
        if (!done1) {
          d1n = bprev;
          if (fprev >= fcur) {
            d1n = min_d;
            done1 = true;
          } else if (fprev < thres && fcur >= thres) {
            done1 = true;
          }
        }
        if (i == 500) bnw1 = bprev;
        bprev = bcur;
        fprev = fcur;
      }
      if (!done1) d1n = bnw1;
      double d2n = (fcur >= thres) ? bcur : (d2cross >= 0.0 ? d2cross : (have_b1 ? b1 : -1.0));
      if (d1n < -1.e-30 || d2n < -1.e-30 || d1n > d2n) {
        set_err(cst, E_ADAPTIVE);
        return;
      }
      d1w = d1n;
      d2w = d2n;
    } else {
      d1w = h.d_1 + 1.0e-8;
      d2w = h.d_2;
    }
    Grid g;
    g.init(d1w, d2w, nbin);
    double bprev = g.next(), fprev = dist_value(h.dist, dp, bprev);
    for (int i = 0; i < nbin; i++) {
      double bcur = g.next();
      double fcur = dist_value(h.dist, dp, bcur);
      double dc = (bprev + bcur) * 0.5, dw = bcur - bprev;
      od[i] = dc;
      odd[i] = dw;
      ond[i] = fprev;
      mass_tot = mass_tot + h.a_ms * (fprev + fcur) / 2.0 * dw * pow(dc, h.b_ms);
      bprev = bcur;
      fprev = fcur;
    }
    f_last = fprev;
  }
  // pass B: mass rescale and trapezoid bin numbers (make_dist.f90:262-283)
  bool resc = s.conserve && (h.moment_in == 3 || h.moment_in == 23 || h.moment_in == 13);
  double ntot = 0.0;
  double fl = resc ? f_last * q_h / mass_tot : f_last;
  for (int i = 0; i < nbin; i++) {
    double fi = resc ? ond[i] * q_h / mass_tot : ond[i];
    double fn = (i + 1 < nbin) ? (resc ? ond[i + 1] * q_h / mass_tot : ond[i + 1]) : fl;
    double n = (fi + fn) / 2.0 * odd[i];
    ond[i] = n;
    ntot += n;
  }
  bool any = false, neg = false;
  double maxd = -1e300;
  for (int i = 0; i < nbin; i++) {
    if (ond[i] < ntot / nbin * 1e-60) ond[i] = 0.0;
    if (ond[i] > 0) any = true;
    if (ond[i] < 0.) neg = true;
    if (od[i] > maxd) maxd = od[i];
  }
  if (maxd <= 0. || neg) {
    set_err(cst, E_MOMENT);
    return;
  }
  // densities and scattering diameters (drop_size_dist.f90:196-204,
  // make_soft_spheroid.f90:78-145); ndens = n / delta (scatProperties.f90:223)
  const double smin = s.softsphere_min;
  for (int i = 0; i < nbin; i++) {
    double dd = od[i];
    double rho, deff;
    if (h.liq_ice == 1) {
      rho = 1000.0;
      deff = dd;
    } else {
      double mass = h.a_ms * pow(dd, h.b_ms);
      double ar = h.as_ratio;
      rho = 0.0;
      deff = 0.0;
      if (h.rho_ms > 0.) {
        rho = h.rho_ms;
        if (ar <= 1.) {
          if (ar < 0.) deff = pow((6.0 * mass) / (PI * h.rho_ms), 1.0 / 3.0);
          if (ar > 0.) deff = pow((6.0 * mass) / (PI * h.rho_ms * ar), 1.0 / 3.0);
        }
        if (ar > 1.) deff = pow((6.0 * mass * pow(ar, 2.0)) / (PI * h.rho_ms), 1.0 / 3.0);
      } else {
        deff = dd;
        if (ar <= 1.) {
          if (ar == -20. || ar == -30.) {
            set_err(cst, E_SOFT);
            return;
          }
          if (ar == -99.) rho = (6.0 * mass) / (PI * pow(dd, 3.0));
          if (ar == -10.) rho = 917.0;
          if (ar > 0.) rho = (6.0 * mass) / (PI * pow(dd, 3.0) * ar);
          if (rho < smin) rho = smin;
          if (ar == -10.) deff = pow((6.0 * mass) / (PI * rho), 1.0 / 3.0);
        }
        if (ar > 1.) rho = (6.0 * mass * pow(ar, 2.0)) / (PI * pow(dd, 3.0));
      }
      if (s.limit_density) {
        if (rho > 917.0)
          rho = 917.0;
        else if (rho < smin)
          rho = smin;
      }
      if (rho <= 0. || deff <= 0.) {
        set_err(cst, E_SOFT);
        return;
      }
    }
    od[i] = deff;
    // the self-similar Rayleigh-Gans model takes the particle mass mass_ds
    // (drop_size_dist.f90:162-165) instead of a density
    orho[i] = (hd.scat_name == PAM_SCAT_SSRG_RT3 && h.liq_ice != 1) ? h.a_ms * pow(dd, h.b_ms) : rho;
    ond[i] = ond[i] / odd[i];
  }
  for (int i = nbin; i < d.nbmax; i++) {
    od[i] = 1.0;
    odd[i] = 1.0;
    ond[i] = 0.0;
    orho[i] = 1000.0;
  }
  *flag = (any && hd.scat_name != PAM_SCAT_DISABLED) ? nbin : 0;
}

void launch_dsd(const Dims &d, const Settings &s, const Profile &p, const pam_hydro_desc *hyd,
                const Hydro4D &h4, Work &w, hipStream_t st) {
  size_t n = (size_t)d.nitems * d.nh;
  if (n == 0) return;
  hipLaunchKernelGGL(k_dsd, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, d, s, p, hyd, h4,
                     w);
}

}  // namespace pam
