// pam_dsd.hip -- drop-size distributions on the device, one wavefront per
// (scattering layer, hydrometeor) with the lanes over the size bins:
// mass-size law (make_mass_size.f90:30-148), distribution parameters
// (make_dist_param.f90:2-631), size grid and bin numbers (make_dist.f90:2-290)
// and soft-sphere density (make_soft_spheroid.f90:78-152).  Each lane finds
// its bin edges in the merged linear/log grid by a merge-path search, so the
// edges and the 500-point adaptive scan are computed in parallel, and the bin
// arrays are stored once, coalesced.
#include "pam_internal.h"

namespace pam {

struct HP {
  double as_ratio, rho_ms, a_ms, b_ms, p_1, p_2, p_3, p_4, d_1, d_2;
  int liq_ice, moment_in, nbin, dist;
};

struct DP {
  double n_0, lambda, mu, gam, n_t, sig, d_ln, d_mono, d_m, n_0_star;
  // the normalised gammas' point-independent factors (make_dist.f90:171-196)
  double b, t1, g1, g2;
};

// the monodisperse families (make_dist_param.f90:133-224); they take the
// three-edge grid of make_dist.f90:88-99
__device__ __forceinline__ bool is_mono(int dist) {
  return dist == PAM_DIST_MONO || dist == PAM_DIST_MONO_COSMO_ICE ||
         dist == PAM_DIST_MONO_ECHAM_CL || dist == PAM_DIST_MONO_ECHAM_ICE;
}

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

// make_dist_params (make_dist_param.f90:2-631).  layer_p and sfc_type
// (land > 0.5) enter the ECHAM monodisperse forms only; 'const' and
// 'const_cosmo_ice' are refused on the host (make_dist never fills their bins).
__device__ int32_t dist_params(const HP &h, double q_h, double n_tot, double r_eff,
                               double layer_t, double layer_p, int sfc_type, DP &o) {
  const double NaN = __builtin_nan("");
  o.n_0 = -99.;
  o.lambda = -99.;
  o.gam = NaN;
  o.mu = NaN;
  o.n_t = -99.;
  o.sig = -99.;
  o.d_ln = NaN;
  o.d_mono = -99.;
  o.d_m = -99.;
  o.n_0_star = -99.;
  o.b = h.b_ms;
  o.t1 = o.g1 = o.g2 = 0.0;
  int mi = h.moment_in;
  if ((mi == 3 && q_h <= 0.) || (mi == 1 && n_tot <= 0.) || (mi == 2 && r_eff <= 0.) ||
      (mi == 13 && (q_h <= 0. || n_tot <= 0.)) || (mi == 23 && (q_h <= 0. || r_eff <= 0.)) ||
      (mi == 12 && (n_tot <= 0. || r_eff <= 0.)))
    return E_DIST_PARAMS;
  const double a = h.a_ms, b = h.b_ms, p1 = h.p_1, p2 = h.p_2, p3 = h.p_3, p4 = h.p_4;
  switch (h.dist) {
    case PAM_DIST_MONO_COSMO_ICE:
    case PAM_DIST_MONO_ECHAM_CL:
    case PAM_DIST_MONO_ECHAM_ICE:
    case PAM_DIST_MONO: {
      o.lambda = 0.0;
      o.mu = 0.0;
      o.gam = 0.0;
      if (h.dist == PAM_DIST_MONO_COSMO_ICE) {
        if (mi == 3 || mi == 13) {
          // N_tot(T) of COSMO's ice, D at most 200 um (make_dist_param.f90:163-178)
          double w1 = 1.e2 * exp(0.2 * (273.15 - layer_t));
          o.n_0 = w1 / DELTA_D_MONO;
          o.d_mono = pow(q_h / (w1 * a), 1.0 / b);
          if (o.d_mono > 2.e-4) {
            o.d_mono = 2.e-4;
            o.n_0 = q_h / (a * pow(o.d_mono, b)) / DELTA_D_MONO;
          }
        }
      } else if (h.dist == PAM_DIST_MONO_ECHAM_CL) {
        // ECHAM6 liquid cloud, land/sea N scaled above 800 hPa
        // (make_dist_param.f90:180-200); kappa is a default-real literal
        double kappa;
        if (sfc_type > 0.5) {
          o.n_0 = 220.e6;
          kappa = sp(1.143);
        } else {
          o.n_0 = 80.e6;
          kappa = sp(1.077);
        }
        if (layer_p < 8.e4) {
          double r = 8.e4 / fmax(1.e4, layer_p);
          o.n_0 = (50.0 + (o.n_0 / 1.e6 - 50.0) * exp(1 - r * r)) * 1.e6;
        }
        const double rho_dry = layer_p / (layer_t * 28704.e-2);  // rho_air.f90
        o.d_mono = kappa * 2.0 * pow(q_h * rho_dry * 3.0 / (4.0 * PI * o.n_0 * 1000.0), 1.0 / 3.0);
        if (o.d_mono > 48.e-6) {
          o.d_mono = 48.e-6;
          o.n_0 = 6. * q_h * rho_dry / (PI * (o.d_mono * o.d_mono * o.d_mono));
        }
      } else if (h.dist == PAM_DIST_MONO_ECHAM_ICE) {
        // ECHAM6 mean ice-crystal volume diameter (make_dist_param.f90:202-209)
        const double rho_dry = layer_p / (layer_t * 28704.e-2);
        double w = 83.8 * pow(1.e3 * q_h * rho_dry, 0.216);
        o.d_mono = 2.e-6 * pow(sqrt(2809.0 * pow(w, 3.0) + 5113188.0) - 2261.0, 1.0 / 3.0);
        o.n_0 = q_h * rho_dry / (DELTA_D_MONO * h.rho_ms * (PI / 6.0) * pow(o.d_mono, 3.0));
      } else {
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
      }
      return (o.n_0 <= 0.0 || o.d_mono <= 0.0) ? E_DIST_PARAMS : 0;
    }
    case PAM_DIST_EXP:
    case PAM_DIST_EXP_FIELD_T:
    case PAM_DIST_EXP_COSMO_SNOW:
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
      if (h.dist == PAM_DIST_EXP_COSMO_SNOW && mi == 3) {
        // COSMO two-moment snow (make_dist_param.f90:343-374): Field et al.
        // 2005's moment relation at n = 3 gives n_0(T, q); its coefficients
        // are default-real literals
        const float mma[10] = {5.065339f, -0.062659f, -3.032362f, 0.029469f, -0.000285f,
                               0.312550f, 0.000204f,  0.003199f,  0.000000f, -0.015952f};
        const float mmb[10] = {0.476221f, -0.015896f, 0.165977f, 0.007468f, -0.000141f,
                               0.060366f, 0.000079f,  0.000594f, 0.000000f, -0.003577f};
        double ztc = layer_t - sp(273.15);
        ztc = fmax(fmin(ztc, 0.0), -40.0);
        const double nn = 3.0, z2 = ztc * ztc;
        double hlp = (double)mma[0] + (double)mma[1] * ztc + (double)mma[2] * nn +
                     (double)mma[3] * ztc * nn + (double)mma[4] * z2 + (double)mma[5] * 9.0 +
                     (double)mma[6] * z2 * nn + (double)mma[7] * ztc * 9.0 +
                     (double)mma[8] * (z2 * ztc) + (double)mma[9] * 27.0;
        const double alf = pow(10.0, hlp);
        const double bet = (double)mmb[0] + (double)mmb[1] * ztc + (double)mmb[2] * nn +
                           (double)mmb[3] * ztc * nn + (double)mmb[4] * z2 + (double)mmb[5] * 9.0 +
                           (double)mmb[6] * z2 * nn + (double)mmb[7] * ztc * 9.0 +
                           (double)mmb[8] * (z2 * ztc) + (double)mmb[9] * 27.0;
        const double m2s = q_h / a;
        const double m3s = alf * exp(bet * log(m2s));
        hlp = 7.628e6 * exp(-0.107 * ztc);
        const double m2s2 = m2s * m2s;
        double n0 = 13.5 * (m2s2 * m2s2) / (m3s * m3s * m3s);
        n0 = fmax(n0, 0.5 * hlp);
        n0 = fmin(n0, 1e2 * hlp);
        n0 = fmin(n0, 1e9);
        n0 = fmax(n0, 1e6);
        o.n_0 = n0;
        o.lambda = pow(a * o.n_0 * g / q_h, 1.0 / (b + 1.0));
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
    case PAM_DIST_MGAMMA_MNH:
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
      if (h.dist == PAM_DIST_MGAMMA_MNH) {
        // Meso-NH: N_tot = C lambda^x, C = p_1, x = p_2 (make_dist_param.f90:486-497)
        if (p1 < 0. || p2 == -99.) return E_DIST_PARAMS;
        double w2 = tgamma((mu + b + 1.0) / gam), w3 = tgamma((mu + 1.0) / gam);
        o.lambda = pow(a * p1 / q_h * w2 / w3, gam / (b - p2));
        o.n_0 = gam * p1 / w3 * pow(o.lambda, (mu + 1.0 + p2) / gam);
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
    case PAM_DIST_NORM_GAMMA: {
      // normalised gamma (make_dist_param.f90:537-600): D_m = M_(b+1)/M_b and
      // N_0* of Testud et al. 2001 with the mass-size exponent b
      if (mi != 0 && mi != 23) return E_DIST_PARAMS;
      if (mi == 0) {
        if (p1 == -99. || p2 == -99. || !((p3 != -99.) ^ (p4 != -99.))) return E_DIST_PARAMS;
        o.d_m = p1;
        o.n_0_star = p2;
        if (p3 != -99.) o.mu = p3;
        if (p4 != -99.) o.mu = p4 - (b + 1);
      } else {
        o.d_m = r_eff * 2;
        const double w1 = tgamma(b + 1);
        const double w2 = w1 / pow(b + 1, b + 1);
        o.n_0_star = q_h / (a * pow(o.d_m, b + 1) * w2);
        if (p3 != -99.) {
          o.mu = p3;
        } else {
          if (h.liq_ice != 1) return E_DIST_PARAMS;
          // Mroz et al. 2023 eq. 21 (rain); the exponent is a default-real literal
          o.mu = 10. * pow(o.d_m * 1000., sp(-0.8)) - 4.;
        }
      }
      const double mu = o.mu;
      o.t1 = tgamma(b + 1) / pow(b + 1, b + 1) * pow(b + mu + 1, b + mu + 1) / tgamma(b + mu + 1);
      return 0;
    }
    case PAM_DIST_NORM_MGAMMA: {
      // normalised modified gamma (make_dist_param.f90:602-624)
      if (p1 == -99. || p2 == -99. || p3 == -99. || p4 == -99. || mi != 0) return E_DIST_PARAMS;
      o.d_m = p1;
      o.n_0_star = p2;
      o.mu = p3;
      o.gam = p4;
      const double mu = p3, gam = p4;
      o.g1 = tgamma((mu + b + 1.) / gam);
      o.g2 = tgamma((mu + b + 2.) / gam);
      o.t1 = gam * tgamma(b + 1) / pow(b + 1, b + 1) * pow(o.g2, b + 1 + mu) / pow(o.g1, b + 2 + mu);
      return 0;
    }
    default:
      return E_DIST_PARAMS;
  }
}


__device__ __forceinline__ double dist_value(int dist, const DP &p, double d) {
  if (dist == PAM_DIST_LOGN)
    return p.n_t / (d * p.sig * sqrt(2.0 * PI)) *
           exp(-(pow(log(d) - p.d_ln, 2.0)) / (2.0 * (p.sig * p.sig)));
  if (dist == PAM_DIST_NORM_GAMMA) {
    const double x = d / p.d_m;
    return p.n_0_star * p.t1 * pow(x, p.mu) * exp(-(p.b + p.mu + 1) * x);
  }
  if (dist == PAM_DIST_NORM_MGAMMA) {
    const double x = d / p.d_m;
    return p.n_0_star * p.t1 * pow(x, p.mu) * exp(-pow(x * p.g2 / p.g1, p.gam));
  }
  return p.n_0 * pow(d, p.mu) * exp(-p.lambda * pow(d, p.gam));
}

// The merged linear/log bin-edge grid of make_dist.f90:120-155: nw+1
// ascending edges on [d1, d2], the union of nlin+1 linear points
// d1 + w_lin (l+1) and nlog+1 log points exp(w_log g + log d1), the log point
// first on ties.  at(i) is the i-th edge without the sequential merge: log
// point g sits at position g + #{l : lin(l) < log(g)}, which grows with g,
// so a binary search over g finds how many log points precede position i.
struct Grid {
  double d1, w_lin, w_log, logd1;
  int nlin, nlog;
  __device__ void init(double a, double b, int nw) {
    nlin = (int)(nw / 2.0);
    nlog = (int)(nw / 2.0);
    if (nlin + nlog == nw) nlin = nlin - 1;
    d1 = a;
    w_lin = (b - a) / (nlin + 1);
    logd1 = log(a);
    w_log = (log(b) - logd1) / (nlog + 1);
  }
  __device__ double lin(int l) const { return d1 + w_lin * (l + 1); }
  __device__ double lg(int g) const { return exp(w_log * g + logd1); }
  // number of linear points strictly below v: an estimate, then corrected
  // against the points themselves (at most nlin + 1 steps, none for NaN)
  __device__ int lin_below(double v) const {
    double est = (v - d1) / w_lin - 1.0;
    int l = !(est > 0.0) ? 0 : (est > (double)(nlin + 1) ? nlin + 1 : (int)est);
    while (l <= nlin && lin(l) < v) l++;
    while (l > 0 && !(lin(l - 1) < v)) l--;
    return l;
  }
  __device__ int pos_log(int g) const { return g + lin_below(lg(g)); }
  __device__ double at(int i) const {
    int lo = -1, hi = nlog;  // largest g with pos_log(g) <= i
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (pos_log(mid) <= i)
        lo = mid;
      else
        hi = mid - 1;
    }
    if (lo >= 0 && pos_log(lo) == i) return lg(lo);
    return lin(i - (lo + 1));
  }
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_imax(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_imin(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}

constexpr int DSD_SCAN = 500;  // points of the adaptive scan (make_dist.f90:119)

// One 64-lane workgroup per (item, hydrometeor) of the chunk.  LDS holds the
// edges and the distribution at the edges: lds_n = max(nbmax, DSD_SCAN) + 1
// doubles each.  Every branch below is uniform across the wavefront.
__global__ void __launch_bounds__(64) k_dsd(Dims d, Settings s, Profile p,
                                            const pam_hydro_desc *hyd, Hydro4D h4, Work w,
                                            int lds_n) {
  extern __shared__ double s_edge[];
  double *sb = s_edge, *sf = s_edge + lds_n;
  const int lane = threadIdx.x;
  const int li = blockIdx.x / d.nh, hh = blockIdx.x % d.nh;
  const int col = w.item_col[li], z = w.item_layer[li];
  const size_t oy = (size_t)col * d.ML + z;
  const uint32_t mask = w.hmask[oy];
  int32_t *flag = w.dsd_flag + (size_t)li * d.nh + hh;
  if (!(mask & (1u << hh))) {
    if (lane == 0) *flag = 0;
    return;
  }
  int32_t *cst = w.col_status + col;
  const pam_hydro_desc &hd = hyd[hh];
  const size_t i4 = oy * d.nh + hh;
  const size_t ob = ((size_t)li * d.nh + hh) * d.nbmax;
  double *od = w.dsd_d + ob, *odd = w.dsd_dd + ob, *ond = w.dsd_nd + ob, *orho = w.dsd_rho + ob;
  // as_ratio of the layer for the self-similar Rayleigh-Gans size along the
  // beam (scatProperties.f90:276, 286-287: negative or NaN reads as 0)
  const double ar4 = h4val(h4.as_ratio, hd.as_ratio, i4);
  if (lane == 0) w.dsd_ar[(size_t)li * d.nh + hh] = (ar4 < 0.0 || isnan(ar4)) ? 0.0 : ar4;
  if (h4.fs_nbin > 0) {
    // full spectrum given (hydrometeor_extinction.f90:120-150): the bins
    // replace the size distribution; the scattering diameter and density
    // are d_ds and rho_ds, delta_d the edge differences, ndens n / delta_d
    const int nb = h4.fs_nbin;
    const size_t b5 = i4 * (size_t)nb, b6 = i4 * (size_t)(nb + 1);
    const bool ssrg = hd.scat_name == PAM_SCAT_SSRG_RT3;
    for (int i = lane; i < d.nbmax; i += 64) {
      if (i < nb) {
        const double dd = h4.fs_d_bound[b6 + i + 1] - h4.fs_d_bound[b6 + i];
        od[i] = h4.fs_d[b5 + i];
        odd[i] = dd;
        ond[i] = h4.fs_n[b5 + i] / dd;
        orho[i] = ssrg ? h4.fs_mass[b5 + i] : h4.fs_rho[b5 + i];
      } else {
        od[i] = 1.0;
        odd[i] = 1.0;
        ond[i] = 0.0;
        orho[i] = 1000.0;
      }
    }
    if (lane == 0) *flag = (hd.scat_name != PAM_SCAT_DISABLED) ? nb : 0;
    return;
  }
  if (lane == 0) *flag = 0;
  HP h;
  h.as_ratio = ar4;
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
  // hydrometeor_extinction.f90:200-202
  if (h.dist == PAM_DIST_MONO || h.dist == PAM_DIST_MONO_COSMO_ICE) h.nbin = 2;
  const bool mono = is_mono(h.dist);
  if (h.nbin > d.nbmax || h.nbin < 1 || (mono && h.nbin != 2)) {
    if (lane == 0) set_err(cst, E_DIST_PARAMS);
    return;
  }
  // absolute moments (q2abs, conversions.f90:5-24)
  const size_t b4 = oy * d.nh;
  double qsum = 0.0;
  bool qnan = false;
  for (int k = 0; k < d.nh; k++) {
    double q = p.hydro_q ? p.hydro_q[b4 + k] : __builtin_nan("");
    if (isnan(q)) qnan = true;
    qsum += q;
  }
  const double T = w.temp[oy], P = w.press[oy], qv = w.q_hum[oy];
  const double qall = (s.incl_hydro_rho && !qnan) ? qsum : 0.0;
  const double den = R_D * (1.0 + (R_V / R_D - 1.0) * qv - qall) * T;
  const double qin = p.hydro_q ? p.hydro_q[b4 + hh] : __builtin_nan("");
  const double nin = p.hydro_n ? p.hydro_n[b4 + hh] : __builtin_nan("");
  const double q_h = qin * P / den, n_tot = nin * P / den;
  const double r_eff = p.hydro_reff ? p.hydro_reff[b4 + hh] : __builtin_nan("");

  int32_t e = mass_size(h);
  DP dp;
  if (!e) e = dist_params(h, q_h, n_tot, r_eff, T, P, p.sfc_type ? p.sfc_type[col] : -9999, dp);
  if (e) {
    if (lane == 0) set_err(cst, e);
    return;
  }
  const int nbin = h.nbin;
  const double thres = 0.1, min_d = sp(1.0e-7), max_d = sp(2.0e-2);
  // edges and the distribution at the edges into LDS
  if (mono) {
    if (lane < 3) {
      const double bv = dp.d_mono + DELTA_D_MONO * 0.5 * (lane - 1);
      const double bnd = lane == 1 ? dp.d_mono : bv;
      sb[lane] = bnd;
      sf[lane] = dist_value(h.dist, dp, bnd);
    }
  } else {
    double d1w, d2w;
    if (s.adaptive) {
      // adaptive scan (make_dist.f90:214-259) over 501 points
      Grid g;
      g.init(min_d, max_d, DSD_SCAN);
      for (int i = lane; i <= DSD_SCAN; i += 64) {
        const double bv = g.at(i);
        sb[i] = bv;
        sf[i] = dist_value(h.dist, dp, bv);
      }
      __syncthreads();
      // search_loop1: the last i with f(i-1) >= thres > f(i); search_loop2:
      // the first i with f(i-1) >= f(i) or f(i-1) < thres <= f(i)
      int last1 = 0, first2 = DSD_SCAN + 1;
      for (int i = lane + 1; i <= DSD_SCAN; i += 64) {
        const double f0 = sf[i - 1], f1 = sf[i];
        if (f0 >= thres && f1 < thres) last1 = i;
        if (first2 > DSD_SCAN && (f0 >= f1 || (f0 < thres && f1 >= thres))) first2 = i;
      }
      last1 = wave_imax(last1);
      first2 = wave_imin(first2);
      const double d2n = (sf[DSD_SCAN] >= thres) ? sb[DSD_SCAN] : (last1 > 0 ? sb[last1] : sb[1]);
      double d1n;
      if (first2 <= DSD_SCAN)
        d1n = (sf[first2 - 1] >= sf[first2]) ? min_d : sb[first2 - 1];
      else
        d1n = sb[DSD_SCAN - 1];
      __syncthreads();  // the scan's LDS is reused below
      if (d1n < -1.e-30 || d2n < -1.e-30 || d1n > d2n) {
        if (lane == 0) set_err(cst, E_ADAPTIVE);
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
    for (int i = lane; i <= nbin; i += 64) {
      const double bv = g.at(i);
      sb[i] = bv;
      sf[i] = dist_value(h.dist, dp, bv);
    }
  }
  __syncthreads();
  // mass of the trapezoid spectrum, rescale and bin numbers (make_dist.f90:262-283)
  double m = 0.0;
  for (int i = lane; i < nbin; i += 64) {
    const double dc = (sb[i] + sb[i + 1]) * 0.5, dw = sb[i + 1] - sb[i];
    m += h.a_ms * (sf[i] + sf[i + 1]) / 2.0 * dw * pow(dc, h.b_ms);
  }
  const double mass_tot = wave_sum(m);
  const bool resc = s.conserve && (h.moment_in == 3 || h.moment_in == 23 || h.moment_in == 13);
  double nsum = 0.0;
  for (int i = lane; i < nbin; i += 64) {
    const double f0 = resc ? sf[i] * q_h / mass_tot : sf[i];
    const double f1 = resc ? sf[i + 1] * q_h / mass_tot : sf[i + 1];
    nsum += (f0 + f1) / 2.0 * (sb[i + 1] - sb[i]);
  }
  const double ntot = wave_sum(nsum);
  // calc_moment's checks (drop_size_dist.f90:176-189): a positive largest
  // diameter and no negative bin
  double maxd = -1e300;
  bool any = false, neg = false;
  for (int i = lane; i < nbin; i += 64) {
    const double f0 = resc ? sf[i] * q_h / mass_tot : sf[i];
    const double f1 = resc ? sf[i + 1] * q_h / mass_tot : sf[i + 1];
    double n = (f0 + f1) / 2.0 * (sb[i + 1] - sb[i]);
    if (n < ntot / nbin * 1e-60) n = 0.0;
    any = any || n > 0;
    neg = neg || n < 0.;
    maxd = fmax(maxd, (sb[i] + sb[i + 1]) * 0.5);
  }
  maxd = wave_max(maxd);
  any = __any(any);
  neg = __any(neg);
  if (maxd <= 0. || neg) {
    if (lane == 0) set_err(cst, E_MOMENT);
    return;
  }
  // densities and scattering diameters (drop_size_dist.f90:196-204,
  // make_soft_spheroid.f90:78-145); ndens = n / delta (scatProperties.f90:223)
  const double smin = s.softsphere_min;
  const bool ssrg = hd.scat_name == PAM_SCAT_SSRG_RT3 && h.liq_ice != 1;
  bool bad = false;
  for (int i = lane; i < d.nbmax; i += 64) {
    if (i >= nbin) {
      od[i] = 1.0;
      odd[i] = 1.0;
      ond[i] = 0.0;
      orho[i] = 1000.0;
      continue;
    }
    const double dd = (sb[i] + sb[i + 1]) * 0.5, dw = sb[i + 1] - sb[i];
    const double f0 = resc ? sf[i] * q_h / mass_tot : sf[i];
    const double f1 = resc ? sf[i + 1] * q_h / mass_tot : sf[i + 1];
    double n = (f0 + f1) / 2.0 * dw;
    if (n < ntot / nbin * 1e-60) n = 0.0;
    double rho, deff;
    if (h.liq_ice == 1) {
      rho = 1000.0;
      deff = dd;
    } else {
      const double mass = h.a_ms * pow(dd, h.b_ms);
      const double ar = h.as_ratio;
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
          if (ar == -20. || ar == -30.) bad = true;  // frequency-dependent density
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
      if (rho <= 0. || deff <= 0.) bad = true;
    }
    od[i] = deff;
    // the self-similar Rayleigh-Gans model takes the particle mass mass_ds
    // (drop_size_dist.f90:162-165) instead of a density
    orho[i] = ssrg ? h.a_ms * pow(dd, h.b_ms) : rho;
    odd[i] = dw;
    ond[i] = n / dw;
  }
  if (__any(bad)) {
    if (lane == 0) set_err(cst, E_SOFT);
    return;
  }
  if (lane == 0) *flag = (any && hd.scat_name != PAM_SCAT_DISABLED) ? nbin : 0;
}

void launch_dsd(const Dims &d, const Settings &s, const Profile &p, const pam_hydro_desc *hyd,
                const Hydro4D &h4, Work &w, hipStream_t st) {
  const size_t n = (size_t)d.nitems * d.nh;
  if (n == 0) return;
  const int lds_n = (d.nbmax > DSD_SCAN ? d.nbmax : DSD_SCAN) + 1;
  hipLaunchKernelGGL(k_dsd, dim3((unsigned)n), dim3(64), 2 * lds_n * sizeof(double), st, d, s, p,
                     hyd, h4, w, lds_n);
}

}  // namespace pam
