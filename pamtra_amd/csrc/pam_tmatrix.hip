// pam_tmatrix.hip -- the T-matrix producer for spheroids: calc_tmatrix and
// calc_single_tmatrix (tmatrix.f90:14-684) over the extended-boundary-
// condition code of tmatrix_amplq.lp.f, for passive runs (every incidence
// angle), one particle orientation (alpha 0, the descriptor's canting angle
// as beta), 30 scattering azimuths.
//
// One workgroup per (layer, channel) unit takes the unit's T-matrix
// hydrometeors and their size bins in the reference's order; per bin it
//   1. solves the T matrix (tmatrix_amplq.lp.f:276-540): NMAX, then NGAUSS,
//      grow until the m = 0 extinction and scattering change by at most
//      1e-3.  The Gauss sums of Q and RG(Q) and the Bessel functions are
//      FP64, Q is rounded to single precision (cmplx without KIND in TT,
//      tmatrix_amplq.lp.f:2130) and inverted in double-double (the
//      reference inverts in binary128), T = -RG(Q) Q^-1 is FP64 and is
//      stored in single precision (COMMON /TMAT/ is REAL*4);
//   2. forms the amplitude matrices (tmatrix_amplq.lp.f:602-888) of the
//      NMU x 2 NMU (incidence, scattering) angle pairs.  With a canting angle
//      the frames are tilted per azimuth (tm_azimuths_beta); with canting 0 the
//      particle frame is the laboratory frame up to the reference's 1e-8
//      angle offsets, so the azimuth enters only through cos(m dphi),
//      sin(m dphi) and the frame rotations: for each m the sums over n, n'
//      are an incidence-side product and a scattering-side product, kept
//      per (pair, m), and the 30 azimuths are evaluated afterwards;
//   3. integrates |S|^2 over the azimuths into the phase matrix, takes the
//      extinction matrix from the forward amplitude and the emission vector
//      as extinction less the integrated scattering (tmatrix.f90:588-672);
//   4. adds part * ndens * del_d to the hydrometeor's sums (tmatrix.f90:
//      423-426), the hydrometeor's sums to the unit's (scatProperties.f90:
//      925-927).
// The unit's sums go to its T slot (downward half, quadrants 1 and 4) and
// R slot (upward half, quadrants 2 and 3), in the precision the doubling
// reads, and to the per-angle extinction and emission arrays.
//
// The Gauss nodes depend only on their number: the host tabulates them
// once per context with the reference's Newton iteration in extended
// precision (tm_gauss_table).
#include "pam_internal.h"

#include <cmath>
#include <vector>

#pragma clang fp contract(off)

namespace pam {

constexpr int TM_WG = 64;
constexpr int TM_NDIR = 2 * NMU;         // scattering directions (hemisphere, angle)
constexpr int TM_NPAIR = NMU * TM_NDIR;  // (incidence angle, scattering direction)
constexpr int TM_NAZ = 30;               // azimuth_num (tmatrix.f90:142)
constexpr int TM_GRP = 2;                // pairs per group of the azimuth stage
constexpr int TM_B = TM_NCAP + 1;        // leading dimension over orders
constexpr int TM_X = 2 * TM_GCAP + 1;    // 1-based point arrays
constexpr int TM_QN = 2 * TM_NCAP;       // largest Q order
constexpr int TM_NACC = TM_NSMAT + 6 * NMU;

// ---- per-workgroup scratch layout (doubles) ----
constexpr size_t TM_O_X = 0, TM_O_W = TM_X, TM_O_S = 2 * TM_X, TM_O_SS = 3 * TM_X,
                 TM_O_R = 4 * TM_X, TM_O_DR = 5 * TM_X, TM_O_DDR = 6 * TM_X, TM_O_DRR = 7 * TM_X,
                 TM_O_DRI = 8 * TM_X;
constexpr size_t TM_BPLANE = (size_t)(TM_GCAP + 1) * TM_B;
constexpr size_t TM_O_BES = 9 * TM_X;                   // 8 planes: j, y, jr, ji, dj, dy, djr, dji
constexpr size_t TM_O_D1 = TM_O_BES + 8 * TM_BPLANE;    // d1, d2 planes
constexpr size_t TM_O_ANN = TM_O_D1 + 2 * TM_BPLANE;    // [TM_B][TM_B]
constexpr size_t TM_QPL = (size_t)TM_QN * TM_QN;
constexpr size_t TM_O_Q = TM_O_ANN + (size_t)TM_B * TM_B;  // 4 planes: re hi, re lo, im hi, im lo
constexpr size_t TM_O_RGQ = TM_O_Q + 4 * TM_QPL;            // 2 planes
constexpr size_t TM_O_TT = TM_O_RGQ + 2 * TM_QPL;           // TR1, TI1
constexpr size_t TM_O_TM = TM_O_TT + 2 * TM_QPL;            // float [8][NCAP][NCAP]
// per (pair, m) amplitude sums at canting 0, or the incidence forms of every
// order [m][NMU][TM_B][8] with a canting angle
constexpr size_t TM_AMP_SIZE = (size_t)TM_NPAIR * TM_B * 8 > (size_t)TM_B * NMU * TM_B * 8
                                   ? (size_t)TM_NPAIR * TM_B * 8
                                   : (size_t)TM_B * NMU * TM_B * 8;
constexpr size_t TM_O_AMP = TM_O_TM + 4 * (size_t)TM_NCAP * TM_NCAP;
constexpr size_t TM_O_E = TM_O_AMP + TM_AMP_SIZE;                       // [NMU][TM_B][8]
constexpr size_t TM_O_DV = TM_O_E + (size_t)NMU * TM_B * 8;     // d functions [2][NDIR][TM_B]
// with a canting angle: the frames of the incidence directions [NMU][6] and of
// the scattering directions [NDIR][NAZ][6] (phi_p, R or R1^-1, cos theta_p),
// and the per-azimuth contributions [NPAIR][NAZ][4]
constexpr size_t TM_O_FR = TM_O_DV + 2 * (size_t)TM_NDIR * TM_B;
constexpr size_t TM_O_CON = TM_O_FR + 6 * (size_t)(NMU + TM_NDIR * TM_NAZ);
constexpr size_t TM_O_PART = TM_O_CON + 4 * (size_t)TM_NPAIR * TM_NAZ;
constexpr size_t TM_O_HYD = TM_O_PART + TM_NACC;
constexpr size_t TM_O_UNIT = TM_O_HYD + TM_NACC;
constexpr size_t TM_SLOT = (TM_O_UNIT + TM_NACC + 63) / 64 * 64;
size_t tm_slot_doubles() { return TM_SLOT; }

// ---- per-call geometry table (doubles), built by k_tm_geom ----
// inc[NMU][5]: phi_p, R11, R12, R21, R22 of the incidence directions;
// out[TM_NDIR][TM_NAZ][5]: phi_p, and the inverse of R1, of the scattering
// directions; dcos[TM_NDIR]: cos theta_p, the argument of the d functions
// (one per direction: independent of the azimuth at canting 0);
// rn[TM_B][TM_B]: sqrt((2n+1)(2n'+1) / (n n' (n+1)(n'+1))).
constexpr size_t TM_G_INC = 0, TM_G_OUT = 5 * NMU, TM_G_COS = TM_G_OUT + 5 * TM_NDIR * TM_NAZ,
                 TM_G_RN = TM_G_COS + TM_NDIR;
constexpr size_t TM_GEO = TM_G_RN + (size_t)TM_B * TM_B;
size_t tm_geo_doubles() { return TM_GEO; }

// ---------------------------------------------------------------------
// double-double arithmetic (error-free transformations with FMA)
// ---------------------------------------------------------------------
struct dd {
  double hi, lo;
};
__device__ __forceinline__ dd two_sum(double a, double b) {
  double s = a + b, bb = s - a;
  return dd{s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ dd fast_sum(double a, double b) {
  double s = a + b;
  return dd{s, b - (s - a)};
}
__device__ __forceinline__ dd dd_add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi), t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = fast_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_sum(s.hi, s.lo);
}
__device__ __forceinline__ dd dd_neg(dd a) { return dd{-a.hi, -a.lo}; }
__device__ __forceinline__ dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }
__device__ __forceinline__ dd dd_mul(dd a, dd b) {
  double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  e = fma(a.hi, b.lo, e);
  e = fma(a.lo, b.hi, e);
  return fast_sum(p, e);
}
__device__ __forceinline__ dd dd_mul_d(dd a, double b) {
  double p = a.hi * b;
  double e = fma(a.hi, b, -p);
  e = fma(a.lo, b, e);
  return fast_sum(p, e);
}
__device__ __forceinline__ dd dd_div(dd a, dd b) {
  double q1 = a.hi / b.hi;
  dd r = dd_sub(a, dd_mul_d(b, q1));
  double q2 = r.hi / b.hi;
  r = dd_sub(r, dd_mul_d(b, q2));
  double q3 = r.hi / b.hi;
  return dd_add(fast_sum(q1, q2), dd{q3, 0.0});
}
struct ddc {
  dd re, im;
};
__device__ __forceinline__ ddc ddc_mul(ddc a, ddc b) {
  return ddc{dd_sub(dd_mul(a.re, b.re), dd_mul(a.im, b.im)),
             dd_add(dd_mul(a.re, b.im), dd_mul(a.im, b.re))};
}
__device__ __forceinline__ ddc ddc_inv(ddc a) {
  dd den = dd_add(dd_mul(a.re, a.re), dd_mul(a.im, a.im));
  return ddc{dd_div(a.re, den), dd_neg(dd_div(a.im, den))};
}

// ---------------------------------------------------------------------
// shared memory of the workgroup (phases use disjoint members)
// ---------------------------------------------------------------------
struct TmLds {
  union {
    struct {  // Gauss sums: per-point weights of TMATR
      double rr[TM_GCAP + 1], ds[TM_GCAP + 1], dss[TM_GCAP + 1];
    } g;
    struct {  // inversion: the pivot column and the pivot's inverse
      double fac[TM_QN][4];
      double inv[4];
      double red[TM_WG / 64];
      int redi[TM_WG / 64];
      int perm[TM_QN];
    } j;
    struct {  // azimuth stage of one group of pairs
      double amp[TM_GRP][TM_B][8];
      double c[TM_GRP][TM_NAZ][4];
    } z;
  } u;
  double red[TM_WG / 64];
  double kext;
  int slot, ok, conv, nnmax1, nnmax2;
};

__device__ __forceinline__ double tm_block_max(TmLds &L, double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) L.red[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = L.red[0];
  for (int i = 1; i < TM_WG / 64; i++) m = fmax(m, L.red[i]);
  __syncthreads();
  return m;
}

// ---------------------------------------------------------------------
// the T-matrix solve (tmatrix_amplq.lp.f), cooperative over the workgroup
// ---------------------------------------------------------------------

// CONST (tmatrix_amplq.lp.f:972-1023), NP = -1, nodes from the table.
// sqrt(float(2n+1)/float(n(n+1))) is a single-precision expression there.
__device__ void tm_const(double *S, const double *gtab, int ngauss, int nmax) {
  const int t = threadIdx.x, ng = 2 * ngauss;
  const double *gx = gtab + (size_t)ngauss * (ngauss - 1) * 2;
  #pragma unroll 1
  for (int i = t; i < ng; i += TM_WG) {
    S[TM_O_X + i + 1] = gx[i];
    S[TM_O_W + i + 1] = gx[ng + i];
  }
  #pragma unroll 1
  for (int k = t; k < TM_B * TM_B; k += TM_WG) {
    const int n = k / TM_B, n1 = k % TM_B;
    if (n < 1 || n1 < 1 || n > nmax || n1 > nmax) continue;
    const float d = sqrtf((float)(2 * n + 1) / (float)(n * (n + 1)));
    const float d1 = sqrtf((float)(2 * n1 + 1) / (float)(n1 * (n1 + 1)));
    S[TM_O_ANN + k] = (double)d * (double)d1 * 0.5;
  }
  __syncthreads();
  #pragma unroll 1
  for (int i = t + 1; i <= ngauss; i += TM_WG) {
    double y = S[TM_O_X + i];
    y = 1.0 / (1.0 - y * y);
    S[TM_O_SS + i] = y;
    S[TM_O_SS + ng - i + 1] = y;
    y = sqrt(y);
    S[TM_O_S + i] = y;
    S[TM_O_S + ng - i + 1] = y;
  }
  __syncthreads();
}

// RJB, RYB, CJB (tmatrix_amplq.lp.f:1383-1520) of one point into the
// Bessel planes, row i; the downward recurrences keep orders <= nmax only.
__device__ void tm_bessel(double *S, int i, double x, double xr, double xi, int nmax, int nnmax1,
                          int nnmax2) {
  double *bj = S + TM_O_BES + (size_t)i * TM_B, *by = bj + TM_BPLANE, *bjr = by + TM_BPLANE,
         *bji = bjr + TM_BPLANE, *bdj = bji + TM_BPLANE, *bdy = bdj + TM_BPLANE,
         *bdjr = bdy + TM_BPLANE, *bdji = bdjr + TM_BPLANE;
  {  // RJB: ratios z(l) downward, kept in bj[1..nmax]
    const int l = nmax + nnmax1;
    const double xx = 1.0 / x;
    double z = 1.0 / ((double)(2 * l + 1) * xx);
    if (l <= nmax) bj[l] = z;
    for (int i1 = l - 1; i1 >= 1; i1--) {
      z = 1.0 / ((double)(2 * i1 + 1) * xx - z);
      if (i1 <= nmax) bj[i1] = z;
    }
    const double z0 = 1.0 / (xx - bj[1]);
    const double y0 = z0 * cos(x) * xx;
    const double y1 = y0 * bj[1];
    bdj[1] = y0 - y1 * xx;
    bj[1] = y1;
    for (int k = 2; k <= nmax; k++) {
      const double yi1 = bj[k - 1], yi = yi1 * bj[k];
      bdj[k] = yi1 - (double)k * yi * xx;
      bj[k] = yi;
    }
  }
  {  // RYB
    const double c = cos(x), s = sin(x);
    const double x1 = 1.0 / x, x2 = x1 * x1, x3 = x2 * x1;
    const double y1 = -c * x2 - s * x1;
    by[1] = y1;
    by[2] = (-3.0 * x3 + x1) * c - 3.0 * x2 * s;
    for (int k = 2; k <= nmax - 1; k++) by[k + 1] = (double)(2 * k + 1) * x1 * by[k] - by[k - 1];
    bdy[1] = -x1 * (c + y1);
    for (int k = 2; k <= nmax; k++) bdy[k] = by[k - 1] - (double)k * x1 * by[k];
  }
  {  // CJB: ratios cz(l) downward, kept in (bjr, bji)[1..nmax]
    const int l = nmax + nnmax2;
    const double xrxi = 1.0 / (xr * xr + xi * xi);
    const double cxxr = xr * xrxi, cxxi = -xi * xrxi;
    double qf = 1.0 / (double)(2 * l + 1);
    double czr = xr * qf, czi = xi * qf;
    if (l <= nmax) {
      bjr[l] = czr;
      bji[l] = czi;
    }
    for (int i1 = l - 1; i1 >= 1; i1--) {
      qf = (double)(2 * i1 + 1);
      const double ar = qf * cxxr - czr, ai = qf * cxxi - czi;
      const double ari = 1.0 / (ar * ar + ai * ai);
      czr = ar * ari;
      czi = -ai * ari;
      if (i1 <= nmax) {
        bjr[i1] = czr;
        bji[i1] = czi;
      }
    }
    double ar = cxxr - bjr[1], ai = cxxi - bji[1];
    const double ari = 1.0 / (ar * ar + ai * ai);
    const double cz0r = ar * ari, cz0i = -ai * ari;
    const double cr = cos(xr) * cosh(xi), ci = -sin(xr) * sinh(xi);
    ar = cz0r * cr - cz0i * ci;
    ai = cz0i * cr + cz0r * ci;
    const double cy0r = ar * cxxr - ai * cxxi, cy0i = ai * cxxr + ar * cxxi;
    const double cy1r = cy0r * bjr[1] - cy0i * bji[1], cy1i = cy0i * bjr[1] + cy0r * bji[1];
    ar = cy1r * cxxr - cy1i * cxxi;
    ai = cy1i * cxxr + cy1r * cxxi;
    bjr[1] = cy1r;
    bji[1] = cy1i;
    bdjr[1] = cy0r - ar;
    bdji[1] = cy0i - ai;
    for (int k = 2; k <= nmax; k++) {
      const double c1r = bjr[k - 1], c1i = bji[k - 1];
      const double zr_ = bjr[k], zi_ = bji[k];
      const double cr_ = c1r * zr_ - c1i * zi_, ci_ = c1i * zr_ + c1r * zi_;
      ar = cr_ * cxxr - ci_ * cxxi;
      ai = ci_ * cxxr + cr_ * cxxi;
      bjr[k] = cr_;
      bji[k] = ci_;
      bdjr[k] = c1r - (double)k * ar;
      bdji[k] = c1i - (double)k * ai;
    }
  }
}

struct TmPart {
  double lam, mrr, mri, a, eps, p;
  double ppi, pir, pii;
};

// VARY (tmatrix_amplq.lp.f:1114-1180) with RSP1 (:1194-1225).  Returns
// false (uniformly) when the Bessel recurrences would exceed the reference's
// array bound.
__device__ bool tm_vary(double *S, TmLds &L, TmPart &P, int ngauss, int nmax) {
  const int t = threadIdx.x, ng = 2 * ngauss;
  {
    const double a = P.a * pow(P.eps, 1.0 / 3.0);
    const double aa = a * a, ee = P.eps * P.eps, ee1 = ee - 1.0;
    #pragma unroll 1
    for (int i = t + 1; i <= ngauss; i += TM_WG) {
      const double c = S[TM_O_X + i], cc = c * c, ss = 1.0 - cc, s = sqrt(ss);
      const double rr = 1.0 / (ss + ee * cc);
      const double r = aa * rr, dr = rr * c * s * ee1;
      S[TM_O_R + i] = r;
      S[TM_O_R + ng - i + 1] = r;
      S[TM_O_DR + i] = dr;
      S[TM_O_DR + ng - i + 1] = -dr;
    }
  }
  const double pi = P.p * 2.0 / P.lam;
  P.ppi = pi * pi;
  P.pir = P.ppi * P.mrr;
  P.pii = P.ppi * P.mri;
  const double v0 = 1.0 / (P.mrr * P.mrr + P.mri * P.mri);
  const double prr = P.mrr * v0, pri = -P.mri * v0;
  __syncthreads();
  double ta = 0.0;
  #pragma unroll 1
  for (int i = t + 1; i <= ng; i += TM_WG) {
    double vv = sqrt(S[TM_O_R + i]);
    const double v = vv * pi;
    ta = fmax(ta, v);
    vv = 1.0 / v;
    S[TM_O_DDR + i] = vv;
    S[TM_O_DRR + i] = prr * vv;
    S[TM_O_DRI + i] = pri * vv;
  }
  ta = tm_block_max(L, ta);
  double tb = ta * sqrt(P.mrr * P.mrr + P.mri * P.mri);
  const double fn = (double)(float)nmax;
  if (fn > tb) tb = fn;
  const int nnmax1 = (int)(8.0 * sqrt(ta > fn ? ta : fn) + 3.0);
  int nnmax2 = (int)(tb + 4.0 * pow(tb, 0.33333) + 8.0 * sqrt(tb));
  nnmax2 = nnmax2 - nmax + 5;
  if (nmax + nnmax2 >= 1200) return false;
  #pragma unroll 1
  for (int i = t + 1; i <= ngauss; i += TM_WG) {
    const double v = sqrt(S[TM_O_R + i]) * pi;
    tm_bessel(S, i, v, v * P.mrr, v * P.mri, nmax, nnmax1, nnmax2);
  }
  __syncthreads();
  return true;
}

// VIG (tmatrix_amplq.lp.f:2049-2097); sqrt(float(i2-1)/float(i2)) is
// single precision there
__device__ void tm_vig(double x, int nmax, int m, double *dv1, double *dv2, int ld) {
  const double qs = sqrt(1.0 - x * x), qs1 = 1.0 / qs;
  if (m == 0) {
    double d1 = 1.0, d2 = x;
    for (int n = 1; n <= nmax; n++) {
      const double qn = n, qn1 = n + 1, qn2 = 2 * n + 1;
      const double d3 = (qn2 * x * d2 - qn * d1) / qn1;
      const double der = qs1 * (qn1 * qn / qn2) * (-d1 + d3);
      dv1[n * ld] = d2;
      dv2[n * ld] = der;
      d1 = d2;
      d2 = d3;
    }
    return;
  }
  double a = 1.0;
  const double qmm = (double)(m * m);
  for (int i = 1; i <= m; i++) {
    const int i2 = i * 2;
    a = a * (double)sqrtf((float)(i2 - 1) / (float)i2) * qs;
  }
  for (int n = 1; n < m; n++) dv1[n * ld] = dv2[n * ld] = 0.0;
  double d1 = 0.0, d2 = a;
  for (int n = m; n <= nmax; n++) {
    const double qn = n, qn2 = 2 * n + 1, qn1 = n + 1;
    const double qnm = sqrt(qn * qn - qmm), qnm1 = sqrt(qn1 * qn1 - qmm);
    const double d3 = (qn2 * x * d2 - qnm * d1) / qnm1;
    const double der = qs1 * (-(qn1 * qnm * d1) + qn * qnm1 * d3) / qn2;
    dv1[n * ld] = d2;
    dv2[n * ld] = der;
    d1 = d2;
    d2 = d3;
  }
}

// Gauss-Jordan inversion in place of the complex double-double matrix in the
// Q planes (order n), partial pivoting by |re| + |im| of the leading parts
// (IZAMAX).  Returns false (uniformly) when a pivot vanishes.
__device__ bool tm_invert(double *S, TmLds &L, int n) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  double *qrh = S + TM_O_Q, *qrl = qrh + TM_QPL, *qih = qrl + TM_QPL, *qil = qih + TM_QPL;
  auto ld = [&](int i, int j) {
    const size_t e = (size_t)i * TM_QN + j;
    return ddc{dd{qrh[e], qrl[e]}, dd{qih[e], qil[e]}};
  };
  auto st = [&](int i, int j, ddc v) {
    const size_t e = (size_t)i * TM_QN + j;
    qrh[e] = v.re.hi;
    qrl[e] = v.re.lo;
    qih[e] = v.im.hi;
    qil[e] = v.im.lo;
  };
  for (int k = 0; k < n; k++) {
    // pivot: first row of the largest |re| + |im| in column k
    double best = -1.0;
    int bi = n;
    #pragma unroll 1
    for (int i = k + t; i < n; i += TM_WG) {
      const size_t e = (size_t)i * TM_QN + k;
      const double v = fabs(qrh[e]) + fabs(qih[e]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o);
      const int oi = __shfl_xor(bi, o);
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    if (lane == 0) {
      L.u.j.red[wv] = best;
      L.u.j.redi[wv] = bi;
    }
    __syncthreads();
    if (t == 0) {
      double b = L.u.j.red[0];
      int p = L.u.j.redi[0];
      for (int q = 1; q < TM_WG / 64; q++)
        if (L.u.j.red[q] > b || (L.u.j.red[q] == b && L.u.j.redi[q] < p)) {
          b = L.u.j.red[q];
          p = L.u.j.redi[q];
        }
      L.u.j.perm[k] = (b > 0.0) ? p : -1;
    }
    __syncthreads();
    const int p = L.u.j.perm[k];
    if (p < 0) return false;
    if (p != k)
      #pragma unroll 1
      for (int j = t; j < n; j += TM_WG) {
        const ddc a = ld(k, j), b = ld(p, j);
        st(k, j, b);
        st(p, j, a);
      }
    __syncthreads();
    if (t == 0) {
      const ddc inv = ddc_inv(ld(k, k));
      L.u.j.inv[0] = inv.re.hi;
      L.u.j.inv[1] = inv.re.lo;
      L.u.j.inv[2] = inv.im.hi;
      L.u.j.inv[3] = inv.im.lo;
    }
    #pragma unroll 1
    for (int i = t; i < n; i += TM_WG) {
      const ddc f = ld(i, k);
      L.u.j.fac[i][0] = f.re.hi;
      L.u.j.fac[i][1] = f.re.lo;
      L.u.j.fac[i][2] = f.im.hi;
      L.u.j.fac[i][3] = f.im.lo;
    }
    __syncthreads();
    const ddc inv = ddc{dd{L.u.j.inv[0], L.u.j.inv[1]}, dd{L.u.j.inv[2], L.u.j.inv[3]}};
    // pivot row times the pivot's inverse (its own entry becomes the inverse)
    #pragma unroll 1
    for (int j = t; j < n; j += TM_WG) st(k, j, j == k ? inv : ddc_mul(ld(k, j), inv));
    __syncthreads();
    // the other rows: A(i, j) -= f_i A(k, j), with A(i, k) taken as 0
    #pragma unroll 1
    for (int e = t; e < n * n; e += TM_WG) {
      const int i = e / n, j = e % n;
      if (i == k) continue;
      const ddc f = ddc{dd{L.u.j.fac[i][0], L.u.j.fac[i][1]}, dd{L.u.j.fac[i][2], L.u.j.fac[i][3]}};
      const ddc akj = ld(k, j);
      const ddc aij = (j == k) ? ddc{dd{0.0, 0.0}, dd{0.0, 0.0}} : ld(i, j);
      const ddc pr = ddc_mul(f, akj);
      st(i, j, ddc{dd_sub(aij.re, pr.re), dd_sub(aij.im, pr.im)});
    }
    __syncthreads();
  }
  // undo the row interchanges as column interchanges, last first
  for (int k = n - 1; k >= 0; k--) {
    const int p = L.u.j.perm[k];
    if (p != k)
      #pragma unroll 1
      for (int i = t; i < n; i += TM_WG) {
        const ddc a = ld(i, k), b = ld(i, p);
        st(i, k, b);
        st(i, p, a);
      }
    __syncthreads();
  }
  return true;
}

// TMATR0 / TMATR (tmatrix_amplq.lp.f:1526-2040) with NCHECK = 1 and TT
// (:2107-2150): the Q and RG(Q) matrices of order m, Q rounded to single
// precision and inverted, T = -RG(Q) Q^-1 in the TT planes.
__device__ bool tm_tmatr(double *S, TmLds &L, const TmPart &P, int m, int ngauss, int nmax) {
  const int t = threadIdx.x;
  // d functions at the points (mirrored half, the one the sums read)
  #pragma unroll 1
  for (int i = t + 1; i <= ngauss; i += TM_WG) {
    const int i1 = ngauss + i, i2 = ngauss - i + 1;
    double *d1 = S + TM_O_D1 + (size_t)i2 * TM_B, *d2 = d1 + TM_BPLANE;
    tm_vig(S[TM_O_X + i1], nmax, m, d1, d2, 1);
    for (int n = 1; n <= nmax; n++) {
      const double si = ((n + m) & 1) ? -1.0 : 1.0;
      d1[n] = d1[n] * si;
      d2[n] = -d2[n] * si;
    }
  }
  const double qm = m, qmm = qm * qm;
  #pragma unroll 1
  for (int i = t + 1; i <= ngauss; i += TM_WG) {
    const double wr = S[TM_O_W + i] * S[TM_O_R + i];
    L.u.g.ds[i] = S[TM_O_S + i] * qm * wr;
    L.u.g.dss[i] = S[TM_O_SS + i] * qmm;
    L.u.g.rr[i] = wr;
  }
  __syncthreads();
  const int mm1 = m > 0 ? m : 1, nm = nmax - mm1 + 1, nq = 2 * nm;
  const double *bj = S + TM_O_BES, *by = bj + TM_BPLANE, *bjr = by + TM_BPLANE,
               *bji = bjr + TM_BPLANE, *bdj = bji + TM_BPLANE, *bdy = bdj + TM_BPLANE,
               *bdjr = bdy + TM_BPLANE, *bdji = bdjr + TM_BPLANE;
  const double *d1p = S + TM_O_D1, *d2p = d1p + TM_BPLANE;
  double *qrh = S + TM_O_Q, *qrl = qrh + TM_QPL, *qih = qrl + TM_QPL, *qil = qih + TM_QPL;
  double *rgr = S + TM_O_RGQ, *rgi = rgr + TM_QPL;
  #pragma unroll 1
  for (int it = t; it < nm * nm; it += TM_WG) {
    const int n1 = mm1 + it / nm, n2 = mm1 + it % nm;
    const double an1 = (double)(float)(n1 * (n1 + 1)), an2 = (double)(float)(n2 * (n2 + 1));
    double ar11 = 0, ar12 = 0, ar21 = 0, ar22 = 0, ai11 = 0, ai12 = 0, ai21 = 0, ai22 = 0;
    double gr11 = 0, gr12 = 0, gr21 = 0, gr22 = 0, gi11 = 0, gi12 = 0, gi21 = 0, gi22 = 0;
    const bool even = ((n1 + n2) & 1) == 0;
    if (m > 0 || even) {
#pragma unroll 1
      for (int i = 1; i <= ngauss; i++) {
        const size_t o1 = (size_t)i * TM_B + n1, o2 = (size_t)i * TM_B + n2;
        const double d1n1 = d1p[o1], d2n1 = d2p[o1], d1n2 = d1p[o2], d2n2 = d2p[o2];
        const double a11 = d1n1 * d1n2, a12 = d1n1 * d2n2, a21 = d2n1 * d1n2, a22 = d2n1 * d2n2;
        const double aa1 = a12 + a21;
        const double qj1 = bj[o1], qy1 = by[o1], qjr2 = bjr[o2], qji2 = bji[o2];
        const double qdjr2 = bdjr[o2], qdji2 = bdji[o2], qdj1 = bdj[o1], qdy1 = bdy[o1];
        const double c1r = qjr2 * qj1, c1i = qji2 * qj1;
        const double b1r = c1r - qji2 * qy1, b1i = c1i + qjr2 * qy1;
        const double c2r = qjr2 * qdj1, c2i = qji2 * qdj1;
        const double b2r = c2r - qji2 * qdy1, b2i = c2i + qjr2 * qdy1;
        const double ddri = S[TM_O_DDR + i];
        const double c3r = ddri * c1r, c3i = ddri * c1i, b3r = ddri * b1r, b3i = ddri * b1i;
        const double c4r = qdjr2 * qj1, c4i = qdji2 * qj1;
        const double b4r = c4r - qdji2 * qy1, b4i = c4i + qdjr2 * qy1;
        const double drri = S[TM_O_DRR + i], drii = S[TM_O_DRI + i];
        const double c5r = c1r * drri - c1i * drii, c5i = c1i * drri + c1r * drii;
        const double b5r = b1r * drri - b1i * drii, b5i = b1i * drri + b1r * drii;
        const double uri = S[TM_O_DR + i], rri = L.u.g.rr[i];
        if (m == 0 || even) {
          const double f1 = m == 0 ? rri * a22 : rri * (a11 * L.u.g.dss[i] + a22);
          double f2 = rri * uri * an1 * a12;
          ar12 = ar12 + f1 * b2r + f2 * b3r;
          ai12 = ai12 + f1 * b2i + f2 * b3i;
          gr12 = gr12 + f1 * c2r + f2 * c3r;
          gi12 = gi12 + f1 * c2i + f2 * c3i;
          f2 = rri * uri * an2 * a21;
          ar21 = ar21 + f1 * b4r + f2 * b5r;
          ai21 = ai21 + f1 * b4i + f2 * b5i;
          gr21 = gr21 + f1 * c4r + f2 * c5r;
          gi21 = gi21 + f1 * c4i + f2 * c5i;
        } else {
          const double c6r = qdjr2 * qdj1, c6i = qdji2 * qdj1;
          const double b6r = c6r - qdji2 * qdy1, b6i = c6i + qdjr2 * qdy1;
          const double c7r = c4r * ddri, c7i = c4i * ddri, b7r = b4r * ddri, b7i = b4i * ddri;
          const double c8r = c2r * drri - c2i * drii, c8i = c2i * drri + c2r * drii;
          const double b8r = b2r * drri - b2i * drii, b8i = b2i * drri + b2r * drii;
          const double dsi = L.u.g.ds[i];
          const double e1 = dsi * aa1;
          ar11 = ar11 + e1 * b1r;
          ai11 = ai11 + e1 * b1i;
          gr11 = gr11 + e1 * c1r;
          gi11 = gi11 + e1 * c1i;
          double e2 = dsi * uri * a11;
          const double e3 = e2 * an2;
          e2 = e2 * an1;
          ar22 = ar22 + e1 * b6r + e2 * b7r + e3 * b8r;
          ai22 = ai22 + e1 * b6i + e2 * b7i + e3 * b8i;
          gr22 = gr22 + e1 * c6r + e2 * c7r + e3 * c8r;
          gi22 = gi22 + e1 * c6i + e2 * c7i + e3 * c8i;
        }
      }
    }
    const double an12 = S[TM_O_ANN + (size_t)n1 * TM_B + n2] * 2.0;
    const double r11 = ar11 * an12, r12 = ar12 * an12, r21 = ar21 * an12, r22 = ar22 * an12;
    const double i11 = ai11 * an12, i12 = ai12 * an12, i21 = ai21 * an12, i22 = ai22 * an12;
    const double rg11 = gr11 * an12, rg12 = gr12 * an12, rg21 = gr21 * an12, rg22 = gr22 * an12;
    const double ig11 = gi11 * an12, ig12 = gi12 * an12, ig21 = gi21 * an12, ig22 = gi22 * an12;
    const double tpir = P.pir, tpii = P.pii, tppi = P.ppi;
    const int k1 = n1 - mm1, kk1 = k1 + nm, k2 = n2 - mm1, kk2 = k2 + nm;
    const double tar11 = -r11, tai11 = -i11, tgr11 = -rg11, tgi11 = -ig11;
    const double tar12 = i12, tai12 = -r12, tgr12 = ig12, tgi12 = -rg12;
    const double tar21 = -i21, tai21 = r21, tgr21 = -ig21, tgi21 = rg21;
    const double tar22 = -r22, tai22 = -i22, tgr22 = -rg22, tgi22 = -ig22;
    // Q enters the inversion rounded to single precision
    auto putq = [&](int r, int c, double re, double im) {
      const size_t e = (size_t)r * TM_QN + c;
      qrh[e] = (double)(float)re;
      qrl[e] = 0.0;
      qih[e] = (double)(float)im;
      qil[e] = 0.0;
    };
    auto putg = [&](int r, int c, double re, double im) {
      const size_t e = (size_t)r * TM_QN + c;
      rgr[e] = re;
      rgi[e] = im;
    };
    putq(k1, k2, tpir * tar21 - tpii * tai21 + tppi * tar12, tpir * tai21 + tpii * tar21 + tppi * tai12);
    putg(k1, k2, tpir * tgr21 - tpii * tgi21 + tppi * tgr12, tpir * tgi21 + tpii * tgr21 + tppi * tgi12);
    putq(kk1, kk2, tpir * tar12 - tpii * tai12 + tppi * tar21, tpir * tai12 + tpii * tar12 + tppi * tai21);
    putg(kk1, kk2, tpir * tgr12 - tpii * tgi12 + tppi * tgr21, tpir * tgi12 + tpii * tgr12 + tppi * tgi21);
    if (m == 0) {
      putq(k1, kk2, 0.0, 0.0);
      putg(k1, kk2, 0.0, 0.0);
      putq(kk1, k2, 0.0, 0.0);
      putg(kk1, k2, 0.0, 0.0);
    } else {
      putq(k1, kk2, tpir * tar11 - tpii * tai11 + tppi * tar22, tpir * tai11 + tpii * tar11 + tppi * tai22);
      putg(k1, kk2, tpir * tgr11 - tpii * tgi11 + tppi * tgr22, tpir * tgi11 + tpii * tgr11 + tppi * tgi22);
      putq(kk1, k2, tpir * tar22 - tpii * tai22 + tppi * tar11, tpir * tai22 + tpii * tar22 + tppi * tai11);
      putg(kk1, k2, tpir * tgr22 - tpii * tgi22 + tppi * tgr11, tpir * tgi22 + tpii * tgr22 + tppi * tgi11);
    }
  }
  __syncthreads();
  if (!tm_invert(S, L, nq)) return false;
  // T = -RG(Q) Q^-1, operands rounded to double, summed over k in order
  double *tr1 = S + TM_O_TT, *ti1 = tr1 + TM_QPL;
  #pragma unroll 1
  for (int e = t; e < nq * nq; e += TM_WG) {
    const int i = e / nq, j = e % nq;
    double tr = 0.0, ti = 0.0;
    for (int k = 0; k < nq; k++) {
      const double arr = rgr[(size_t)i * TM_QN + k], ari = rgi[(size_t)i * TM_QN + k];
      const double ar = qrh[(size_t)k * TM_QN + j], ai = qih[(size_t)k * TM_QN + j];
      tr = tr - arr * ar + ari * ai;
      ti = ti - arr * ai - ari * ar;
    }
    tr1[(size_t)i * TM_QN + j] = tr;
    ti1[(size_t)i * TM_QN + j] = ti;
  }
  __syncthreads();
  return true;
}

// T(m) of the TT planes into the single-precision store [c][n-1][n'-1],
// c = RT11, RT12, RT21, RT22, IT11, IT12, IT21, IT22 (tmatrix_amplq.lp.f:
// 468-528)
__device__ void tm_store(double *S, int m, int nmax) {
  const int t = threadIdx.x;
  float *T = reinterpret_cast<float *>(S + TM_O_TM);
  const double *tr1 = S + TM_O_TT, *ti1 = tr1 + TM_QPL;
  const int mm1 = m > 0 ? m : 1, nm = nmax - mm1 + 1;
  #pragma unroll 1
  for (int e = t; e < nm * nm; e += TM_WG) {
    const int a = e / nm, b = e % nm;  // 0-based within the block
    const int n = mm1 + a, nn = mm1 + b;
    const size_t o = (size_t)(n - 1) * TM_NCAP + (nn - 1), pl = (size_t)TM_NCAP * TM_NCAP;
    const size_t e11 = (size_t)a * TM_QN + b, e12 = (size_t)a * TM_QN + b + nm,
                 e21 = (size_t)(a + nm) * TM_QN + b, e22 = (size_t)(a + nm) * TM_QN + b + nm;
    T[0 * pl + o] = (float)tr1[e11];
    T[1 * pl + o] = (float)tr1[e12];
    T[2 * pl + o] = (float)tr1[e21];
    T[3 * pl + o] = (float)tr1[e22];
    T[4 * pl + o] = (float)ti1[e11];
    T[5 * pl + o] = (float)ti1[e12];
    T[6 * pl + o] = (float)ti1[e21];
    T[7 * pl + o] = (float)ti1[e22];
  }
  __syncthreads();
}

// The frame of one direction (tmatrix_amplq.lp.f:630-745) with alpha = 0
// and the canting angle beta [deg] (the reference's 1e-8 offsets included):
// cos theta_p, phi_p and R (incidence) or R1 (scattering).
__device__ void tm_frame(double tl, double pl, double beta, double &ctp, double &phip,
                         double R[2][2]) {
  const double pin = acos(-1.0), pin2 = pin * 0.5, pi = pin / 180.0;
  const double alph = 0.0 * pi, eps = 1e-8;
  double bet = beta * pi;
  double thetl = tl * pi, phil = pl * pi;
  if (thetl < pin2) thetl = thetl + eps;
  if (thetl > pin2) thetl = thetl - eps;
  if (phil < pin) phil = phil + eps;
  if (phil > pin) phil = phil - eps;
  if (bet <= pin2 && pin2 - bet <= eps) bet = bet - eps;
  if (bet > pin2 && bet - pin2 <= eps) bet = bet + eps;
  const double cb = cos(bet), sb = sin(bet), ct = cos(thetl), st = sin(thetl);
  double cp = cos(phil - alph), sp = sin(phil - alph);
  ctp = ct * cb + st * sb * cp;
  const double thetp = acos(ctp);
  const double cpp = cb * st * cp - sb * ct, spp = st * sp;
  phip = atan(spp / cpp);
  if (phip > 0.0 && sp < 0.0) phip = phip + pin;
  if (phip < 0.0 && sp > 0.0) phip = phip + pin;
  if (phip < 0.0) phip = phip + 2.0 * pin;
  const double ca = cos(alph), sa = sin(alph);
  const double B[3][3] = {{ca * cb, sa * cb, -sb}, {-sa, ca, 0.0}, {ca * sb, sa * sb, cb}};
  cp = cos(phil);
  sp = sin(phil);
  const double AL[3][2] = {{ct * cp, -sp}, {ct * sp, cp}, {-st, 0.0}};
  const double ct2 = ctp, st2 = sin(thetp), cp2 = cos(phip), sp2 = sin(phip);
  const double AP[2][3] = {{ct2 * cp2, ct2 * sp2, -st2}, {-sp2, cp2, 0.0}};
  double C[3][2];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 2; j++) {
      double x = 0.0;
      for (int k = 0; k < 3; k++) x = x + B[i][k] * AL[k][j];
      C[i][j] = x;
    }
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      double x = 0.0;
      for (int k = 0; k < 3; k++) x = x + AP[i][k] * C[k][j];
      R[i][j] = x;
    }
}

// VIGAMPL (tmatrix_amplq.lp.f:905-960): d functions / sin and derivative
__device__ void tm_vigampl(double x, int nmax, int m, double *dv1, double *dv2) {
  for (int n = 1; n <= nmax; n++) dv1[n] = dv2[n] = 0.0;
  const double dx = fabs(x);
  if (fabs(1.0 - dx) <= 1e-10) {
    if (m != 1) return;
    for (int n = 1; n <= nmax; n++) {
      double dn = (double)(n * (n + 1));
      dn = 0.5 * sqrt(dn);
      if (x < 0.0) dn = dn * (((n + 1) & 1) ? -1.0 : 1.0);
      dv1[n] = dn;
      if (x < 0.0) dn = -dn;
      dv2[n] = dn;
    }
    return;
  }
  double a = 1.0;
  const double qs = sqrt(1.0 - x * x), qs1 = 1.0 / qs, dsi = qs1;
  if (m == 0) {
    double d1 = 1.0, d2 = x;
    for (int n = 1; n <= nmax; n++) {
      const double qn = n, qn1 = n + 1, qn2 = 2 * n + 1;
      const double d3 = (qn2 * x * d2 - qn * d1) / qn1;
      const double der = qs1 * (qn1 * qn / qn2) * (-d1 + d3);
      dv1[n] = d2 * dsi;
      dv2[n] = der;
      d1 = d2;
      d2 = d3;
    }
    return;
  }
  const double qmm = (double)(m * m);
  for (int i = 1; i <= m; i++) {
    const int i2 = i * 2;
    a = a * sqrt((double)(i2 - 1) / (double)i2) * qs;
  }
  double d1 = 0.0, d2 = a;
  for (int n = m; n <= nmax; n++) {
    const double qn = n, qn2 = 2 * n + 1, qn1 = n + 1;
    const double qnm = sqrt(qn * qn - qmm), qnm1 = sqrt(qn1 * qn1 - qmm);
    const double d3 = (qn2 * x * d2 - qnm * d1) / qnm1;
    const double der = qs1 * (-(qn1 * qnm * d1) + qn * qnm1 * d3) / qn2;
    dv1[n] = d2 * dsi;
    dv2[n] = der;
    d1 = d2;
    d2 = d3;
  }
}

// The incidence-side forms of order m (see tm_amp_order) into E[jj][n][8],
// from the stored T(m) and the d functions va = dv1, vb = dv2 [jj][n] of the
// incidence directions.
__device__ void tm_eforms(const double *S, const double *va, const double *vb, const double *rn,
                          int m, int nmax, double *E) {
  const int t = threadIdx.x;
  const int nmin = m > 1 ? m : 1;
  const float *T = reinterpret_cast<const float *>(S + TM_O_TM);
  const size_t pl = (size_t)TM_NCAP * TM_NCAP;
  const double fm = (double)m;
  #pragma unroll 1
  for (int it = t; it < NMU * nmax; it += TM_WG) {
    const int jj = it / nmax, n = 1 + it % nmax;
    double e[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (n >= nmin)
#pragma unroll 1
      for (int nn = nmin; nn <= nmax; nn++) {
        const size_t o = (size_t)(n - 1) * TM_NCAP + (nn - 1);
        const double c = fm * va[jj * TM_B + nn], d = vb[jj * TM_B + nn];
        // CAL = i^(nn - n - 1) rn(n, nn)
        const double r = rn[(size_t)n * TM_B + nn];
        const int q = ((nn - n - 1) % 4 + 4) % 4;
        const double cr = q == 0 ? r : (q == 2 ? -r : 0.0), ci = q == 1 ? r : (q == 3 ? -r : 0.0);
        const double t11r = T[0 * pl + o], t12r = T[1 * pl + o], t21r = T[2 * pl + o],
                     t22r = T[3 * pl + o], t11i = T[4 * pl + o], t12i = T[5 * pl + o],
                     t21i = T[6 * pl + o], t22i = T[7 * pl + o];
        // x1 = c T11 + d T12, x2 = c T21 + d T22, x3 = d T11 + c T12, x4 = d T21 + c T22
        const double x1r = c * t11r + d * t12r, x1i = c * t11i + d * t12i;
        const double x2r = c * t21r + d * t22r, x2i = c * t21i + d * t22i;
        const double x3r = d * t11r + c * t12r, x3i = d * t11i + c * t12i;
        const double x4r = d * t21r + c * t22r, x4i = d * t21i + c * t22i;
        e[0] += cr * x1r - ci * x1i;
        e[1] += cr * x1i + ci * x1r;
        e[2] += cr * x2r - ci * x2i;
        e[3] += cr * x2i + ci * x2r;
        e[4] += cr * x3r - ci * x3i;
        e[5] += cr * x3i + ci * x3r;
        e[6] += cr * x4r - ci * x4i;
        e[7] += cr * x4i + ci * x4r;
      }
    double *eo = E + ((size_t)jj * TM_B + n) * 8;
    for (int k = 0; k < 8; k++) eo[k] = e[k];
  }
  __syncthreads();
}

// The amplitude sums of order m (tmatrix_amplq.lp.f:808-855), factored:
//   E1(jj,n) = sum_n' CAL(n,n') (c T11 + d T12), E2 = sum CAL (c T21 + d T22),
//   E3 = sum CAL (d T11 + c T12),               E4 = sum CAL (d T21 + c T22),
// with c = m dv1(n'), d = dv2(n') of the incidence direction jj, then per
// (jj, scattering direction o), with a = m dv1(n), b = dv2(n) of o:
//   A_VV = sum_n a E1 + b E2,  A_VH = sum a E3 + b E4,
//   A_HV = sum b E1 + a E2,    A_HH = sum b E3 + a E4,
// so that VV = sum_m 2cos(m dphi) A_VV (A_VV alone for m = 0),
// VH = sum 2sin A_VH, HV = -sum 2sin A_HV, HH = sum 2cos A_HH.
__device__ void tm_amp_order(double *S, TmLds &L, const double *geo, int m, int nmax) {
  const int t = threadIdx.x;
  const int nmin = m > 1 ? m : 1;
  // d functions of every direction: a = dv1, b = dv2 [dir][n]
  double *va = S + TM_O_DV, *vb = va + (size_t)TM_NDIR * TM_B;
#pragma unroll 1
  for (int dir = t; dir < TM_NDIR; dir += TM_WG)
    tm_vigampl(geo[TM_G_COS + dir], nmax, m, va + (size_t)dir * TM_B, vb + (size_t)dir * TM_B);
  __syncthreads();
  double *E = S + TM_O_E;
  tm_eforms(S, va, vb, geo + TM_G_RN, m, nmax, E);
  const double fm = (double)m;
  double *A = S + TM_O_AMP;
  #pragma unroll 1
  for (int pr = t; pr < TM_NPAIR; pr += TM_WG) {
    const int jj = pr / TM_NDIR, o = pr % TM_NDIR;
    double vv[2] = {0, 0}, vh[2] = {0, 0}, hv[2] = {0, 0}, hh[2] = {0, 0};
#pragma unroll 1
    for (int n = nmin; n <= nmax; n++) {
      const double a = fm * va[o * TM_B + n], b = vb[o * TM_B + n];
      const double *e = E + ((size_t)jj * TM_B + n) * 8;
      vv[0] += a * e[0] + b * e[2];
      vv[1] += a * e[1] + b * e[3];
      vh[0] += a * e[4] + b * e[6];
      vh[1] += a * e[5] + b * e[7];
      hv[0] += b * e[0] + a * e[2];
      hv[1] += b * e[1] + a * e[3];
      hh[0] += b * e[4] + a * e[6];
      hh[1] += b * e[5] + a * e[7];
    }
    double *ao = A + ((size_t)pr * TM_B + m) * 8;
    ao[0] = vv[0];
    ao[1] = vv[1];
    ao[2] = vh[0];
    ao[3] = vh[1];
    ao[4] = hv[0];
    ao[5] = hv[1];
    ao[6] = hh[0];
    ao[7] = hh[1];
  }
  __syncthreads();
}

// One (pair, azimuth) of the azimuth sums (tmatrix_amplq.lp.f:857-869,
// tmatrix.f90:616-639): the amplitude matrix from the order sums (before the
// division by k), the frame rotations R (gi[1..4]) and R1^-1 (go[1..4]),
// |S|^2 into the four contributions c, and the imaginary parts of the
// amplitudes S11, S22 (the forward extinction).
__device__ __forceinline__ void tm_item(double vv[2], double vh[2], double hv[2], double hh[2],
                                        const double *gi, const double *go, double dk,
                                        double wave_num, double fact_sca, double phi_w,
                                        double c[4], double &im11, double &im22) {
  for (int q = 0; q < 2; q++) {
    vv[q] = vv[q] / dk;
    vh[q] = vh[q] / dk;
    hv[q] = hv[q] / dk;
    hh[q] = hh[q] / dk;
  }
  // rotations R (incidence) and R1^-1 (scattering), tmatrix_amplq.lp.f:862-869
  const double r11 = gi[1], r12 = gi[2], r21 = gi[3], r22 = gi[4];
  const double s11 = go[1], s12 = go[2], s21 = go[3], s22 = go[4];
  double cvv[2], cvh[2], chv[2], chh[2], o11[2], o12[2], o21[2], o22[2];
  for (int q = 0; q < 2; q++) {
    cvv[q] = vv[q] * r11 + vh[q] * r21;
    cvh[q] = vv[q] * r12 + vh[q] * r22;
    chv[q] = hv[q] * r11 + hh[q] * r21;
    chh[q] = hv[q] * r12 + hh[q] * r22;
  }
  for (int q = 0; q < 2; q++) {
    o11[q] = (s11 * cvv[q] + s12 * chv[q]) * wave_num;
    o12[q] = (s11 * cvh[q] + s12 * chh[q]) * wave_num;
    o21[q] = (s21 * cvv[q] + s22 * chv[q]) * wave_num;
    o22[q] = (s21 * cvh[q] + s22 * chh[q]) * wave_num;
  }
  const double a11 = o11[0] * o11[0] - o11[1] * (-o11[1]);
  const double a12 = o12[0] * o12[0] - o12[1] * (-o12[1]);
  const double a21 = o21[0] * o21[0] - o21[1] * (-o21[1]);
  const double a22 = o22[0] * o22[0] - o22[1] * (-o22[1]);
  c[0] = (fact_sca * (((a11 + a12) + a21) + a22)) * phi_w;
  c[1] = (fact_sca * (((a11 - a12) + a21) - a22)) * phi_w;
  c[2] = (fact_sca * (((a11 + a12) - a21) - a22)) * phi_w;
  c[3] = (fact_sca * (((a11 - a12) - a21) + a22)) * phi_w;
  im11 = o11[1];
  im22 = o22[1];
}

// The azimuth sums (tmatrix.f90:588-663) of every pair from the per-order
// amplitude sums, then the emission vector (:664-672).  part: scattering
// matrix (2,NMU,2,NMU,2), extinction (2,2,NMU), emission (2,NMU), Fortran
// order, of one particle.
__device__ void tm_azimuths(double *S, TmLds &L, const double *geo, int nmax, double lam,
                            const double *mu, const double *qw) {
  const int t = threadIdx.x;
  double *part = S + TM_O_PART, *pext = part + TM_NSMAT, *pemis = pext + 4 * NMU;
  const double wave_num = 2.0 * PI / lam;
  const double dk = 2.0 * acos(-1.0) / lam;
  const double fact_sca = 0.5 / (wave_num * wave_num);
  const double fext = (2.0 * PI) / pow(wave_num, 2.0);
  const double phi_w = 1.0 / 360.0 * (360.0 / TM_NAZ), phi0_w = 1.0 / 360.0 * (360.0 / 1);
  #pragma unroll 1
  for (int k = t; k < 4 * NMU; k += TM_WG) pext[k] = 0.0;
  const double *A = S + TM_O_AMP;
  for (int g0 = 0; g0 < TM_NPAIR; g0 += TM_GRP) {
    #pragma unroll 1
    for (int k = t; k < TM_GRP * (nmax + 1) * 8; k += TM_WG) {
      const int p = k / ((nmax + 1) * 8), r = k % ((nmax + 1) * 8);
      L.u.z.amp[p][r / 8][r % 8] = A[((size_t)(g0 + p) * TM_B + r / 8) * 8 + r % 8];
    }
    __syncthreads();
    #pragma unroll 1
    for (int it = t; it < TM_GRP * TM_NAZ; it += TM_WG) {
      const int p = it / TM_NAZ, az = it % TM_NAZ;
      const int pr = g0 + p, jj = pr / TM_NDIR, o = pr % TM_NDIR;
      const double *gi = geo + TM_G_INC + 5 * jj;
      const double *go = geo + TM_G_OUT + 5 * ((size_t)o * TM_NAZ + az);
      const double ph = go[0] - gi[0];
      double vv[2] = {L.u.z.amp[p][0][0], L.u.z.amp[p][0][1]}, vh[2] = {0.0, 0.0},
             hv[2] = {0.0, 0.0}, hh[2] = {L.u.z.amp[p][0][6], L.u.z.amp[p][0][7]};
      // cos(m dphi), sin(m dphi) by the rotation recurrence from m = 1
      double s1, c1;
      sincos(ph, &s1, &c1);
      double cs = c1, sn = s1;
#pragma unroll 1
      for (int m = 1; m <= nmax; m++) {
        const double fc = 2.0 * cs, fs = 2.0 * sn;
        const double cn = cs * c1 - sn * s1;
        sn = sn * c1 + cs * s1;
        cs = cn;
        const double *a = L.u.z.amp[p][m];
        vv[0] += fc * a[0];
        vv[1] += fc * a[1];
        vh[0] += fs * a[2];
        vh[1] += fs * a[3];
        hv[0] -= fs * a[4];
        hv[1] -= fs * a[5];
        hh[0] += fc * a[6];
        hh[1] += fc * a[7];
      }
      double im11, im22;
      tm_item(vv, vh, hv, hh, gi, go, dk, wave_num, fact_sca, phi_w, L.u.z.c[p][az], im11, im22);
      if (az == 0 && o == jj) {  // forward: thet0 == thet, phi0 == phi
        const double ip = (im11 + im22) * fext, imn = (im11 - im22) * fext;
        pext[0 + 2 * (0 + 2 * jj)] = 0.0 + phi0_w * ip;
        pext[0 + 2 * (1 + 2 * jj)] = 0.0 + phi0_w * imn;
        pext[1 + 2 * (0 + 2 * jj)] = 0.0 + phi0_w * imn;
        pext[1 + 2 * (1 + 2 * jj)] = 0.0 + phi0_w * ip;
      }
    }
    __syncthreads();
    if (t < TM_GRP) {
      const int pr = g0 + t, jj = pr / TM_NDIR, o = pr % TM_NDIR, kk = o / NMU, ll = o % NMU;
      double s[4] = {0.0, 0.0, 0.0, 0.0};
      for (int az = 0; az < TM_NAZ; az++)
        for (int q = 0; q < 4; q++) s[q] = s[q] + L.u.z.c[t][az][q];
      // smat(a, ll, b, jj, kk)
      auto at = [&](int a, int b) { return part + a + 2 * (ll + NMU * (b + 2 * (jj + NMU * kk))); };
      *at(0, 0) = 0.0 + s[0] * phi0_w;
      *at(0, 1) = 0.0 + s[1] * phi0_w;
      *at(1, 0) = 0.0 + s[2] * phi0_w;
      *at(1, 1) = 0.0 + s[3] * phi0_w;
    }
    __syncthreads();
  }
  if (t < NMU) {
    const int jj = t;
    double e11 = 0.0, e12 = 0.0;
    for (int kk = 0; kk < 2; kk++)
      for (int ll = 0; ll < NMU; ll++) {
        const double z11 = part[0 + 2 * (ll + NMU * (0 + 2 * (jj + NMU * kk)))];
        const double z12 = part[0 + 2 * (ll + NMU * (1 + 2 * (jj + NMU * kk)))];
        e11 = e11 + z11 * qw[ll] * 2.0 * PI;
        e12 = e12 + z12 * qw[ll] * 2.0 * PI;
      }
    pemis[0 + 2 * jj] = pext[0 + 2 * (0 + 2 * jj)] - e11;
    pemis[1 + 2 * jj] = pext[0 + 2 * (1 + 2 * jj)] - e12;
  }
  __syncthreads();
}

// ---- a canting angle: the particle frame turns with the scattering azimuth

// Frames of every incidence direction (azimuth 0) and scattering direction
// for the canting angle beta [deg] (tmatrix.f90:588-612 angles, the frame
// of tmatrix_amplq.lp.f:630-760) into the frame region of the scratch.
__device__ void tm_frames_beta(double *S, double beta, const double *mu) {
  const float thet_max = 179.9999f;
  double *fr = S + TM_O_FR;
#pragma unroll 1
  for (int k = threadIdx.x; k < NMU + TM_NDIR * TM_NAZ; k += TM_WG) {
    double th, phi;
    double *g;
    bool inc = k < NMU;
    if (inc) {
      th = acos(mu[k] * 1.0) * 180.0 / PI;
      phi = 0.0;
      g = fr + 6 * k;
    } else {
      const int q = k - NMU, o = q / TM_NAZ, az = q % TM_NAZ, kk = o / NMU, ll = o % NMU;
      th = acos(mu[ll] * (kk == 0 ? 1.0 : -1.0)) * 180.0 / PI;
      phi = 360.0 / (double)(float)TM_NAZ * ((double)(float)(az + 1) - 1.0);
      g = fr + 6 * ((size_t)NMU + (size_t)o * TM_NAZ + az);
    }
    if (th > (double)thet_max) th = 180.0;
    double ctp, phip, R[2][2];
    tm_frame(th, phi, beta, ctp, phip, R);
    g[0] = phip;
    if (inc) {
      g[1] = R[0][0];
      g[2] = R[0][1];
      g[3] = R[1][0];
      g[4] = R[1][1];
    } else {
      const double dt = 1.0 / (R[0][0] * R[1][1] - R[0][1] * R[1][0]);
      g[1] = R[1][1] * dt;
      g[2] = -R[0][1] * dt;
      g[3] = -R[1][0] * dt;
      g[4] = R[0][0] * dt;
    }
    g[5] = ctp;
  }
  __syncthreads();
}

// The incidence forms of order m at the tilted incidence frames, kept for
// every m: E_all[m][jj][n][8] in the amplitude region.
__device__ void tm_eforms_beta(double *S, const double *geo, int m, int nmax) {
  double *va = S + TM_O_DV, *vb = va + (size_t)TM_NDIR * TM_B;
#pragma unroll 1
  for (int jj = threadIdx.x; jj < NMU; jj += TM_WG)
    tm_vigampl(S[TM_O_FR + 6 * jj + 5], nmax, m, va + (size_t)jj * TM_B, vb + (size_t)jj * TM_B);
  __syncthreads();
  tm_eforms(S, va, vb, geo + TM_G_RN, m, nmax, S + TM_O_AMP + (size_t)m * NMU * TM_B * 8);
}

// The scattering-side sums of order m of one pair: VIGAMPL's recurrence at
// the scattering direction's cos theta_p (tmatrix_amplq.lp.f:905-960), run
// in n, accumulating A_VV, A_VH, A_HV, A_HH (see tm_amp_order) from the
// incidence forms e[n][8] as the d functions come.
__device__ __forceinline__ void tm_vig_sum(double x, int nmax, int m, const double *e,
                                           double acc[8]) {
  for (int k = 0; k < 8; k++) acc[k] = 0.0;
  const int nmin = m > 1 ? m : 1;
  const double fm = (double)m;
  auto add = [&](int n, double dv1, double dv2) {
    if (n < nmin) return;
    const double a = fm * dv1, b = dv2;
    const double *q = e + (size_t)n * 8;
    acc[0] += a * q[0] + b * q[2];
    acc[1] += a * q[1] + b * q[3];
    acc[2] += a * q[4] + b * q[6];
    acc[3] += a * q[5] + b * q[7];
    acc[4] += b * q[0] + a * q[2];
    acc[5] += b * q[1] + a * q[3];
    acc[6] += b * q[4] + a * q[6];
    acc[7] += b * q[5] + a * q[7];
  };
  if (fabs(1.0 - fabs(x)) <= 1e-10) {
    if (m != 1) return;
#pragma unroll 1
    for (int n = 1; n <= nmax; n++) {
      double dn = 0.5 * sqrt((double)(n * (n + 1)));
      if (x < 0.0) dn = dn * (((n + 1) & 1) ? -1.0 : 1.0);
      const double dv1 = dn;
      if (x < 0.0) dn = -dn;
      add(n, dv1, dn);
    }
    return;
  }
  const double qs = sqrt(1.0 - x * x), qs1 = 1.0 / qs, dsi = qs1;
  if (m == 0) {
    double d1 = 1.0, d2 = x;
#pragma unroll 1
    for (int n = 1; n <= nmax; n++) {
      const double qn = n, qn1 = n + 1, qn2 = 2 * n + 1;
      const double d3 = (qn2 * x * d2 - qn * d1) / qn1;
      const double der = qs1 * (qn1 * qn / qn2) * (-d1 + d3);
      add(n, d2 * dsi, der);
      d1 = d2;
      d2 = d3;
    }
    return;
  }
  double a = 1.0;
  const double qmm = (double)(m * m);
#pragma unroll 1
  for (int i = 1; i <= m; i++) {
    const int i2 = i * 2;
    a = a * sqrt((double)(i2 - 1) / (double)i2) * qs;
  }
  double d1 = 0.0, d2 = a;
#pragma unroll 1
  for (int n = m; n <= nmax; n++) {
    const double qn = n, qn2 = 2 * n + 1, qn1 = n + 1;
    const double qnm = sqrt(qn * qn - qmm), qnm1 = sqrt(qn1 * qn1 - qmm);
    const double d3 = (qn2 * x * d2 - qnm * d1) / qnm1;
    const double der = qs1 * (-(qn1 * qnm * d1) + qn * qnm1 * d3) / qn2;
    add(n, d2 * dsi, der);
    d1 = d2;
    d2 = d3;
  }
}

// The azimuth sums with a canting angle: per (pair, azimuth) the sums over m
// of the scattering-side forms at that azimuth's tilted frame, then as
// tm_azimuths.
__device__ void tm_azimuths_beta(double *S, int nmax, double lam, const double *qw) {
  const int t = threadIdx.x;
  double *part = S + TM_O_PART, *pext = part + TM_NSMAT, *pemis = pext + 4 * NMU;
  const double wave_num = 2.0 * PI / lam;
  const double dk = 2.0 * acos(-1.0) / lam;
  const double fact_sca = 0.5 / (wave_num * wave_num);
  const double fext = (2.0 * PI) / pow(wave_num, 2.0);
  const double phi_w = 1.0 / 360.0 * (360.0 / TM_NAZ), phi0_w = 1.0 / 360.0 * (360.0 / 1);
#pragma unroll 1
  for (int k = t; k < 4 * NMU; k += TM_WG) pext[k] = 0.0;
  __syncthreads();
  const double *fr = S + TM_O_FR, *Eall = S + TM_O_AMP;
  double *con = S + TM_O_CON;
#pragma unroll 1
  for (int it = t; it < TM_NPAIR * TM_NAZ; it += TM_WG) {
    const int pr = it / TM_NAZ, az = it % TM_NAZ, jj = pr / TM_NDIR, o = pr % TM_NDIR;
    const double *gi = fr + 6 * jj, *go = fr + 6 * ((size_t)NMU + (size_t)o * TM_NAZ + az);
    const double ph = go[0] - gi[0];
    double s1, c1;
    sincos(ph, &s1, &c1);
    double cs = c1, sn = s1;
    double vv[2] = {0.0, 0.0}, vh[2] = {0.0, 0.0}, hv[2] = {0.0, 0.0}, hh[2] = {0.0, 0.0};
    double acc[8];
#pragma unroll 1
    for (int m = 0; m <= nmax; m++) {
      tm_vig_sum(go[5], nmax, m, Eall + ((size_t)m * NMU + jj) * TM_B * 8, acc);
      if (m == 0) {
        vv[0] += acc[0];
        vv[1] += acc[1];
        hh[0] += acc[6];
        hh[1] += acc[7];
        continue;
      }
      const double fc = 2.0 * cs, fs = 2.0 * sn;
      const double cn = cs * c1 - sn * s1;
      sn = sn * c1 + cs * s1;
      cs = cn;
      vv[0] += fc * acc[0];
      vv[1] += fc * acc[1];
      vh[0] += fs * acc[2];
      vh[1] += fs * acc[3];
      hv[0] -= fs * acc[4];
      hv[1] -= fs * acc[5];
      hh[0] += fc * acc[6];
      hh[1] += fc * acc[7];
    }
    double im11, im22;
    tm_item(vv, vh, hv, hh, gi, go, dk, wave_num, fact_sca, phi_w,
            con + ((size_t)pr * TM_NAZ + az) * 4, im11, im22);
    if (az == 0 && o == jj) {
      const double ip = (im11 + im22) * fext, imn = (im11 - im22) * fext;
      pext[0 + 2 * (0 + 2 * jj)] = 0.0 + phi0_w * ip;
      pext[0 + 2 * (1 + 2 * jj)] = 0.0 + phi0_w * imn;
      pext[1 + 2 * (0 + 2 * jj)] = 0.0 + phi0_w * imn;
      pext[1 + 2 * (1 + 2 * jj)] = 0.0 + phi0_w * ip;
    }
  }
  __syncthreads();
#pragma unroll 1
  for (int pr = t; pr < TM_NPAIR; pr += TM_WG) {
    const int jj = pr / TM_NDIR, o = pr % TM_NDIR, kk = o / NMU, ll = o % NMU;
    double sm[4] = {0.0, 0.0, 0.0, 0.0};
    for (int az = 0; az < TM_NAZ; az++)
      for (int q = 0; q < 4; q++) sm[q] = sm[q] + con[((size_t)pr * TM_NAZ + az) * 4 + q];
    auto at = [&](int a, int b) { return part + a + 2 * (ll + NMU * (b + 2 * (jj + NMU * kk))); };
    *at(0, 0) = 0.0 + sm[0] * phi0_w;
    *at(0, 1) = 0.0 + sm[1] * phi0_w;
    *at(1, 0) = 0.0 + sm[2] * phi0_w;
    *at(1, 1) = 0.0 + sm[3] * phi0_w;
  }
  __syncthreads();
  if (t < NMU) {
    const int jj = t;
    double e11 = 0.0, e12 = 0.0;
    for (int kk = 0; kk < 2; kk++)
      for (int ll = 0; ll < NMU; ll++) {
        const double z11 = part[0 + 2 * (ll + NMU * (0 + 2 * (jj + NMU * kk)))];
        const double z12 = part[0 + 2 * (ll + NMU * (1 + 2 * (jj + NMU * kk)))];
        e11 = e11 + z11 * qw[ll] * 2.0 * PI;
        e12 = e12 + z12 * qw[ll] * 2.0 * PI;
      }
    pemis[0 + 2 * jj] = pext[0 + 2 * (0 + 2 * jj)] - e11;
    pemis[1 + 2 * jj] = pext[0 + 2 * (1 + 2 * jj)] - e12;
  }
  __syncthreads();
}

// One particle: the T matrix and its phase matrix, extinction and emission
// into the part accumulator.  Returns false (uniformly) on a failure the
// reference reports as fatal, or where the T-matrix order or the Gauss
// points would exceed the device limits.
__device__ bool tm_particle(double *S, TmLds &L, const double *geo, const double *gtab,
                            double freq_hz, double mrr, double mri, double axi, double as_ratio,
                            double beta, const double *mu, const double *qw, int *nmax_out) {
  const int t = threadIdx.x;
  TmPart P;
  P.lam = C_LIGHT / freq_hz;
  P.mrr = mrr;
  P.mri = fabs(mri);
  P.eps = 1.0 / as_ratio;
  P.p = acos(-1.0);
  P.a = 1.0 * axi;
  const double ddelt = 0.001;
  const double xev = 2.0 * P.p * P.a / P.lam;
  const int ixxx = (int)(xev + 4.05 * pow(xev, 0.333333));
  const int inm1 = ixxx > 4 ? ixxx : 4;
  if (inm1 > TM_NCAP) return false;
  double qext1 = 0.0, qsca1 = 0.0;
  int nmax = 0, ngauss = 0;
  bool conv = false;
  auto q0 = [&](double &qe, double &qs) {
    const double *tr1 = S + TM_O_TT, *ti1 = tr1 + TM_QPL;
    qe = 0.0;
    qs = 0.0;
    for (int n = 1; n <= nmax; n++) {
      const int n1 = n + nmax;
      const double a = tr1[(size_t)(n - 1) * TM_QN + n - 1], b = ti1[(size_t)(n - 1) * TM_QN + n - 1];
      const double c = tr1[(size_t)(n1 - 1) * TM_QN + n1 - 1], d = ti1[(size_t)(n1 - 1) * TM_QN + n1 - 1];
      const double dn1 = (double)(2 * n + 1);
      qs = qs + dn1 * (a * a + b * b + c * c + d * d);
      qe = qe + (a + c) * dn1;
    }
  };
  for (int nma = inm1; nma <= TM_NCAP; nma++) {
    nmax = nma;
    ngauss = nmax * 2;
    if (ngauss > TM_GCAP) return false;
    tm_const(S, gtab, ngauss, nmax);
    if (!tm_vary(S, L, P, ngauss, nmax)) return false;
    if (!tm_tmatr(S, L, P, 0, ngauss, nmax)) return false;
    double qext, qsca;
    q0(qext, qsca);
    const double dsca = fabs((qsca1 - qsca) / qsca), dext = fabs((qext1 - qext) / qext);
    qext1 = qext;
    qsca1 = qsca;
    if (dsca <= ddelt && dext <= ddelt) {
      conv = true;
      break;
    }
  }
  if (!conv) return false;
  conv = false;
  for (int ngaus = ngauss + 1; ngaus <= TM_GCAP; ngaus++) {
    ngauss = ngaus;
    tm_const(S, gtab, ngauss, nmax);
    if (!tm_vary(S, L, P, ngauss, nmax)) return false;
    if (!tm_tmatr(S, L, P, 0, ngauss, nmax)) return false;
    double qext, qsca;
    q0(qext, qsca);
    const double dsca = fabs((qsca1 - qsca) / qsca), dext = fabs((qext1 - qext) / qext);
    qext1 = qext;
    qsca1 = qsca;
    if (dsca <= ddelt && dext <= ddelt) {
      conv = true;
      break;
    }
  }
  // the reference goes on to 500 points; beyond the device's limit the
  // particle is refused rather than taken unconverged
  if (!conv) return false;
  *nmax_out = nmax;
  // canting 0: the factored amplitude sums; otherwise the tilted frames
  const bool tilt = beta != 0.0;
  if (tilt) tm_frames_beta(S, beta, mu);
  tm_store(S, 0, nmax);
  if (tilt)
    tm_eforms_beta(S, geo, 0, nmax);
  else
    tm_amp_order(S, L, geo, 0, nmax);
  for (int m = 1; m <= nmax; m++) {
    if (!tm_tmatr(S, L, P, m, ngauss, nmax)) return false;
    tm_store(S, m, nmax);
    if (tilt)
      tm_eforms_beta(S, geo, m, nmax);
    else
      tm_amp_order(S, L, geo, m, nmax);
  }
  if (tilt)
    tm_azimuths_beta(S, nmax, P.lam, qw);
  else
    tm_azimuths(S, L, geo, nmax, P.lam, mu, qw);
  (void)t;
  return true;
}

// One (layer, channel) unit: the sums over its T-matrix hydrometeors and
// their bins (tmatrix.f90:150-426, scatProperties.f90:925-927), written to
// the unit's slots and per-angle arrays.
__device__ void tm_unit(const Dims &d, const Settings &s, const pam_hydro_desc *hyd,
                        const Hydro4D &h4, const double *freqs, const Work &w, int itf,
                        double *S, TmLds &L) {
  const int t = threadIdx.x;
  const int li = itf / d.nf, f = itf % d.nf;
  double *part = S + TM_O_PART, *hsum = S + TM_O_HYD, *usum = S + TM_O_UNIT;
  #pragma unroll 1
  for (int k = t; k < TM_NACC; k += TM_WG) usum[k] = 0.0;
  if (t == 0) {
    L.kext = 0.0;
    w.kext[itf] = 0.0;
    w.salb[itf] = 0.0;
    w.back[itf] = 0.0;
    w.nleg[itf] = 0;
  }
  __syncthreads();
  bool ok = true;
  unsigned long long nbins = 0, nsum = 0;
  for (int h = 0; h < d.nh && ok; h++) {
    const pam_hydro_desc &hd = hyd[h];
    const int nbin = w.dsd_flag[(size_t)li * d.nh + h];
    if (hd.scat_name != PAM_SCAT_TMATRIX || nbin <= 0) continue;
    #pragma unroll 1
    for (int k = t; k < TM_NACC; k += TM_WG) hsum[k] = 0.0;
    const double *mi = w.mie_m + (size_t)itf * 4 + (hd.liq_ice == 1 ? 0 : 2);
    const cd ref = cd{mi[0], -mi[1]};  // m = n - i k (scatProperties.f90:243)
    // the descriptor's (or 4D) as_ratio and canting, or with full spectra the
    // per-bin lists; negative or NaN read as 0 (scatProperties.f90:284-287)
    const double ar0 = w.dsd_ar[(size_t)li * d.nh + h];
    const double cant0 = (hd.canting < 0.0 || isnan(hd.canting)) ? 0.0 : hd.canting;
    const bool fs = h4.fs_nbin > 0;
    const size_t b5 =
        fs ? (((size_t)w.item_col[li] * d.ML + w.item_layer[li]) * d.nh + h) * (size_t)h4.fs_nbin
           : 0;
    const size_t ob = ((size_t)li * d.nh + h) * d.nbmax;
    for (int ir = 0; ir < nbin && ok; ir++) {
      const double nd = w.dsd_nd[ob + ir], dd_ = w.dsd_dd[ob + ir];
      const double rho = w.dsd_rho[ob + ir], dmax = w.dsd_d[ob + ir];
      if (nd == 0.0) continue;
      double ar = ar0, cant = cant0;
      if (fs) {
        ar = h4.fs_as_ratio[b5 + ir];
        cant = h4.fs_canting[b5 + ir];
        ar = (ar < 0.0 || isnan(ar)) ? 0.0 : ar;
        cant = (cant < 0.0 || isnan(cant)) ? 0.0 : cant;
      }
      if (!(nd > 0.0) || !(dd_ > 0.0) || !(rho > 0.0) || !(ar > 0.0)) {
        ok = false;
        break;
      }
      const cd mmix = (hd.liq_ice == -1 && rho != 917.0) ? eps_mix_air(ref, rho) : ref;
      // mindex = conjg(mMix)
      const double axi = ar <= 1.0 ? 0.5 * dmax * pow(ar, 1.0 / 3.0)
                                   : 0.5 * dmax / pow(ar, 2.0 / 3.0);
      int nm = 0;
      __syncthreads();
      if (!tm_particle(S, L, w.tm_geo, w.tm_gauss, freqs[f] * 1.e9, mmix.re, -mmix.im, axi, ar,
                       cant, w.mu, w.qw, &nm)) {
        ok = false;
        break;
      }
      nbins++;
      nsum += (unsigned long long)nm;
      #pragma unroll 1
      for (int k = t; k < TM_NACC; k += TM_WG) hsum[k] = hsum[k] + part[k] * nd * dd_;
      __syncthreads();
    }
    if (!ok) break;
    #pragma unroll 1
    for (int k = t; k < TM_NACC; k += TM_WG) usum[k] = usum[k] + hsum[k];
    if (t == 0) L.kext = L.kext + hsum[TM_NSMAT + 0 + 2 * (0 + 2 * (NMU - 1))];
    __syncthreads();
  }
  if (!ok) {
    if (t == 0) set_err(&w.itf_status[itf], E_TMATRIX);
    return;
  }
  // the slots: T (+,+) from the downward half, R (-,+) from the upward half;
  // slot entry (r, c) = smat(r, c) with r = (angle, Stokes) out, c = in
  const bool f32 = phase_in_f32(s, w);
  double *Tq = w.T + (size_t)itf * NN2, *Rq = w.R + (size_t)itf * NN2;
  #pragma unroll 1
  for (int e = t; e < NN2; e += TM_WG) {
    const int r = e / NN, c = e % NN;
    const double tv = usum[r + NN * c], rv = usum[NN2 + r + NN * c];
    if (f32) {
      reinterpret_cast<float *>(Tq)[e] = (float)tv;
      reinterpret_cast<float *>(Rq)[e] = (float)rv;
    } else {
      Tq[e] = tv;
      Rq[e] = rv;
    }
  }
  const double *uext = usum + TM_NSMAT, *uemis = uext + 4 * NMU;
  #pragma unroll 1
  for (int jj = t; jj < NMU; jj += TM_WG) {
    w.tm_ext[((size_t)itf * NMU + jj) * 2 + 0] = uext[0 + 2 * (0 + 2 * jj)];
    w.tm_ext[((size_t)itf * NMU + jj) * 2 + 1] = uext[0 + 2 * (1 + 2 * jj)];
    w.tm_emis[((size_t)itf * NMU + jj) * 2 + 0] = uemis[0 + 2 * jj];
    w.tm_emis[((size_t)itf * NMU + jj) * 2 + 1] = uemis[1 + 2 * jj];
  }
  if (t == 0) {
    w.kext[itf] = L.kext;
    atomicAdd(&w.counters[6], nbins);
    atomicAdd(&w.counters[7], nsum);
  }
}

__global__ void __launch_bounds__(TM_WG) __attribute__((amdgpu_waves_per_eu(2))) k_tmatrix(Dims d, Settings s, const pam_hydro_desc *hyd,
                                                   Hydro4D h4, const double *freqs, Work w) {
  __shared__ TmLds L;
  double *S = w.tm_scr + (size_t)blockIdx.x * TM_SLOT;
  const int n = d.nitems * d.nf;
  for (int itf = next_unit(&w.queue[0], &L.slot); itf < n; itf = next_unit(&w.queue[0], &L.slot))
    tm_unit(d, s, hyd, h4, freqs, w, itf, S, L);
}

// calc_tmatrix (tmatrix.f90:150-426) over the bins of one hydrometeor, per
// bin as_ratio, canting 0, refractive index m = mre - i mim: one workgroup;
// out = the sums, scattering matrix (2,NMU,2,NMU,2), extinction (2,2,NMU),
// emission (2,NMU), Fortran order; status 0 or E_TMATRIX.
__global__ void __launch_bounds__(TM_WG) k_tm_hydro(int nbins, const double *dmax,
                                                    const double *del_d, const double *ndens,
                                                    const double *rho, const double *as_ratio,
                                                    const double *canting, double freq_hz,
                                                    double mre, double mim,
                                                    int phase, const double *geo,
                                                    const double *gtab, const double *mu,
                                                    const double *qw, double *S, double *out,
                                                    int *status, int *nmax_out) {
  __shared__ TmLds L;
  const int t = threadIdx.x;
  double *part = S + TM_O_PART, *hsum = S + TM_O_HYD;
  #pragma unroll 1
  for (int k = t; k < TM_NACC; k += TM_WG) hsum[k] = 0.0;
  __syncthreads();
  const cd ref = cd{mre, -mim};
  int st = 0;
  for (int ir = 0; ir < nbins; ir++) {
    const double nd = ndens[ir], dd_ = del_d[ir], ar = as_ratio[ir], cant = canting[ir];
    if (nd == 0.0) continue;
    if (!(nd > 0.0) || !(dd_ > 0.0) || !(rho[ir] > 0.0) || !(ar > 0.0) || !(cant >= 0.0)) {
      st = E_TMATRIX;
      break;
    }
    const cd mmix = (phase == -1 && rho[ir] != 917.0) ? eps_mix_air(ref, rho[ir]) : ref;
    const double axi = ar <= 1.0 ? 0.5 * dmax[ir] * pow(ar, 1.0 / 3.0)
                                 : 0.5 * dmax[ir] / pow(ar, 2.0 / 3.0);
    int nm = 0;
    __syncthreads();
    if (!tm_particle(S, L, geo, gtab, freq_hz, mmix.re, -mmix.im, axi, ar, cant, mu, qw, &nm)) {
      st = E_TMATRIX;
      break;
    }
    if (t == 0) nmax_out[ir] = nm;
    #pragma unroll 1
    for (int k = t; k < TM_NACC; k += TM_WG) hsum[k] = hsum[k] + part[k] * nd * dd_;
    __syncthreads();
  }
  #pragma unroll 1
  for (int k = t; k < TM_NACC; k += TM_WG) out[k] = hsum[k];
  if (t == 0) *status = st;
}

// Geometry of the call: frames of every incidence and scattering direction
// (tmatrix.f90:588-612: thet0 = acos(mu) in degrees, thet = acos(+-mu),
// 180 above 179.9999 (a single-precision literal), phi = 12 (n - 1)), and
// the CAL magnitudes.
__global__ void k_tm_geom(const double *mu, double *geo) {
  const float thet_max = 179.9999f;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < NMU + TM_NDIR * TM_NAZ + TM_B * TM_B;
       k += gridDim.x * blockDim.x) {
    if (k < NMU) {
      double th = acos(mu[k] * 1.0) * 180.0 / PI;
      if (th > (double)thet_max) th = 180.0;
      double ctp, phip, R[2][2];
      tm_frame(th, 0.0, 0.0, ctp, phip, R);
      double *g = geo + TM_G_INC + 5 * k;
      g[0] = phip;
      g[1] = R[0][0];
      g[2] = R[0][1];
      g[3] = R[1][0];
      g[4] = R[1][1];
    } else if (k < NMU + TM_NDIR * TM_NAZ) {
      const int q = k - NMU, o = q / TM_NAZ, az = q % TM_NAZ, kk = o / NMU, ll = o % NMU;
      double th = acos(mu[ll] * (kk == 0 ? 1.0 : -1.0)) * 180.0 / PI;
      if (th > (double)thet_max) th = 180.0;
      const double phi = 360.0 / (double)(float)TM_NAZ * ((double)(float)(az + 1) - 1.0);
      double ctp, phip, R[2][2];
      tm_frame(th, phi, 0.0, ctp, phip, R);
      const double dt = 1.0 / (R[0][0] * R[1][1] - R[0][1] * R[1][0]);
      double *g = geo + TM_G_OUT + 5 * ((size_t)o * TM_NAZ + az);
      g[0] = phip;
      g[1] = R[1][1] * dt;
      g[2] = -R[0][1] * dt;
      g[3] = -R[1][0] * dt;
      g[4] = R[0][0] * dt;
      if (az == 0) geo[TM_G_COS + o] = ctp;
    } else {
      const int q = k - NMU - TM_NDIR * TM_NAZ, n = q / TM_B, nn = q % TM_B;
      double v = 0.0;
      if (n >= 1 && nn >= 1) {
        double dnn = (double)((2 * n + 1) * (2 * nn + 1));
        dnn = dnn / (double)(n * nn * (n + 1) * (nn + 1));
        v = sqrt(dnn);
      }
      geo[TM_G_RN + q] = v;
    }
  }
}

// Gauss-Legendre points and weights for NG = 2 NGAUSS, NGAUSS = 1..TM_GCAP,
// by QGAUSS's Newton iteration (tmatrix_amplq.lp.f:1025-1092) in extended
// precision, rounded to double: table[g(g-1)*2 ...] = x[1..2g], w[1..2g].
std::vector<double> tm_gauss_table() {
  std::vector<double> tab;
  tab.reserve((size_t)TM_GCAP * (TM_GCAP + 1) * 2);
  typedef long double Q;
  for (int g = 1; g <= TM_GCAP; g++) {
    const int n = 2 * g;
    std::vector<Q> z(n + 1), wt(n + 1);
    const Q A = 1, B = 2, C = 3;
    const int ind = n % 2, k = n / 2 + ind;
    const Q f = n;
    for (int i = 1; i <= k; i++) {
      const int m = n + 1 - i;
      Q x = 0, pa = 0, pb, pc;
      if (i == 1) x = A - B / ((f + A) * f);
      if (i == 2) x = (z[n] - A) * 4 + z[n];
      if (i == 3) x = (z[n - 1] - z[n]) * 1.6L + z[n - 1];
      if (i > 3) x = (z[m + 1] - z[m + 2]) * C + z[m + 3];
      if (i == k && ind == 1) x = 0;
      int niter = 0;
      Q check = 1e-32L;
      for (;;) {
        pb = 1;
        niter++;
        if (niter > 100) check = check * 10;
        pc = x;
        Q dj = A;
        for (int j = 2; j <= n; j++) {
          dj = dj + A;
          pa = pb;
          pb = pc;
          pc = x * pb + (x * pb - pa) * (dj - A) / dj;
        }
        pa = A / ((pb - x * pc) * f);
        pb = pa * pc * (A - x * x);
        x = x - pb;
        if (!(std::fabs(pb) > check * std::fabs(x))) break;
      }
      z[m] = x;
      wt[m] = B * (pa * pa * (A - x * x));
      if (i == k && ind == 1) continue;
      z[i] = -z[m];
      wt[i] = wt[m];
    }
    for (int i = 1; i <= n; i++) tab.push_back((double)z[i]);
    for (int i = 1; i <= n; i++) tab.push_back((double)wt[i]);
  }
  return tab;
}

void launch_tmatrix(const Dims &d, const Settings &s, const pam_hydro_desc *hyd,
                    const Hydro4D &h4, const double *freqs, Work &w, hipStream_t st) {
  const size_t n = (size_t)d.nitems * d.nf;
  if (n == 0) return;
  launch_mie_index(d, hyd, freqs, w, st);
  hipLaunchKernelGGL(k_tm_geom, dim3(8), dim3(256), 0, st, w.mu, w.tm_geo);
  const unsigned grid = (unsigned)std::min<size_t>(n, (size_t)w.tm_slots);
  hipLaunchKernelGGL(k_tmatrix, dim3(grid), dim3(TM_WG), 0, st, d, s, hyd, h4, freqs, w);
}

size_t tm_nacc() { return TM_NACC; }

void launch_tm_hydro(int nbins, const double *dmax, const double *del_d, const double *ndens,
                     const double *rho, const double *as_ratio, const double *canting,
                     double freq_hz, double mre, double mim, int phase, const double *mu,
                     const double *qw, const double *gtab, double *geo, double *scr, double *out,
                     int *status, int *nmax_out, hipStream_t st) {
  hipLaunchKernelGGL(k_tm_geom, dim3(8), dim3(256), 0, st, mu, geo);
  hipLaunchKernelGGL(k_tm_hydro, dim3(1), dim3(TM_WG), 0, st, nbins, dmax, del_d, ndens, rho,
                     as_ratio, canting, freq_hz, mre, mim, phase, geo, gtab, mu, qw, scr, out,
                     status, nmax_out);
}

}  // namespace pam
