/* pam_tmatrix.c -- CPU oracle of the T-matrix producer for spheroids.
 *
 * TEST INFRASTRUCTURE ONLY.  tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it (through libpam_oracle.so) as the
 * checker of the device path; the product never does.
 *
 * Restates calc_tmatrix and calc_single_tmatrix (src/tmatrix.f90:14-684)
 * over the extended-boundary-condition T-matrix code of Mishchenko that the
 * reference vendors (src/tmatrix_amplq.lp.f, src/tmatrix_lpq.f), for the
 * case the reference runs: spheroids (NP = -1), equal-volume radius
 * (RAT = 1), one particle orientation (ALPHA = 0, BETA = canting), one
 * incidence azimuth and 30 scattering azimuths, Lobatto quadrature.
 *
 * Precision follows the reference variable by variable: binary128
 * (__float128, the reference's REAL*16) for the quadrature, the spheroid
 * shape, the Bessel functions, the Q matrices and their inversion; double
 * for the product T = -RG(Q) Q^-1 (TT, REAL*8 implicit typing) and the
 * amplitude matrix; single precision where the reference calls default-kind
 * intrinsics (float() and sqrt of REAL*4 in CONST and VIG) and for the stored
 * T matrix (COMMON /TMAT/ is REAL*4), and the Q matrix enters its inversion
 * rounded to single precision (cmplx without KIND in TT).
 *
 * Pinned by tests/test_tmatrix_oracle.py against the reference's own
 * Fortran (oracle/_ref, built by oracle/ref.mk) and, without it, against
 * the sphere limit (as_ratio = 1 reduces the T matrix to the Mie
 * coefficients: extinction and scattering against the Mie series).
 */
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/pamtra_amd.h"

typedef __float128 qf;

#define NMU PAM_NUMMU
#define NST PAM_NSTOKES

/* array limits of tmatrix_amplq.par.f */
#define TM_NPN1 100
#define TM_NPNG1 500
#define TM_NPNG2 (2 * TM_NPNG1)
#define TM_NPN2 (2 * TM_NPN1)

void pam_oracle_lobatto(int num, double *mu, double *w);
void pam_oracle_eps_mix(double re1, double im1, double mix_de, double *re, double *im);

static const double PI_D = 3.141592653589793; /* constants.f90:16 */
static const double C_LIGHT = 299792458.0;    /* constants.f90:15 */

/* Working storage of one T-matrix solve.  Index conventions are the
 * reference's (1-based): points 1..NG, orders 1..NMAX. */
typedef struct {
  qf x[TM_NPNG2 + 1], w[TM_NPNG2 + 1], s[TM_NPNG2 + 1], ss[TM_NPNG2 + 1];
  qf an[TM_NPN1 + 1], ann[TM_NPN1 + 1][TM_NPN1 + 1];
  qf r[TM_NPNG2 + 1], dr[TM_NPNG2 + 1], ddr[TM_NPNG2 + 1], drr[TM_NPNG2 + 1],
      dri[TM_NPNG2 + 1];
  qf ppi, pir, pii;
  /* Bessel functions [point][order] */
  qf bj[TM_NPNG1 + 1][TM_NPN1 + 1], by[TM_NPNG1 + 1][TM_NPN1 + 1];
  qf bjr[TM_NPNG1 + 1][TM_NPN1 + 1], bji[TM_NPNG1 + 1][TM_NPN1 + 1];
  qf bdj[TM_NPNG1 + 1][TM_NPN1 + 1], bdy[TM_NPNG1 + 1][TM_NPN1 + 1];
  qf bdjr[TM_NPNG1 + 1][TM_NPN1 + 1], bdji[TM_NPNG1 + 1][TM_NPN1 + 1];
  /* Wigner d functions at the points [point][order] */
  qf d1[TM_NPNG2 + 1][TM_NPN1 + 1], d2[TM_NPNG2 + 1][TM_NPN1 + 1];
  /* Q and RG(Q), 0-based [row][col], 2 NM square */
  qf qr[TM_NPN2][TM_NPN2], qi[TM_NPN2][TM_NPN2], rgqr[TM_NPN2][TM_NPN2],
      rgqi[TM_NPN2][TM_NPN2];
  /* T = -RG(Q) Q^-1, 0-based */
  double tr1[TM_NPN2][TM_NPN2], ti1[TM_NPN2][TM_NPN2];
} tm_work;

/* The stored T matrix: 8 REAL*4 components (RT11, RT12, RT21, RT22, IT11,
 * IT12, IT21, IT22) of [m = 0..nmax][n = 1..nmax][n' = 1..nmax]. */
typedef struct {
  int nmax;
  float *t; /* [8][nmax + 1][nmax][nmax] */
} tm_tmat;

static inline float *tm_at(const tm_tmat *T, int c, int m, int n, int nn) {
  int N = T->nmax;
  return &T->t[(((size_t)c * (N + 1) + m) * N + (n - 1)) * N + (nn - 1)];
}

/* QGAUSS (tmatrix_amplq.lp.f:1025-1092) with IND1 = IND2 = 0: Gauss-Legendre
 * points z[1..n] (ascending) and weights by Newton iteration on P_n. */
static void tm_qgauss(int n, qf *z, qf *w) {
  const qf A = 1, B = 2, C = 3;
  int ind = n % 2, k = n / 2 + ind;
  qf f = n;
  for (int i = 1; i <= k; i++) {
    int m = n + 1 - i;
    qf x = 0, pa = 0, pb, pc;
    if (i == 1) x = A - B / ((f + A) * f);
    if (i == 2) x = (z[n] - A) * 4 + z[n];
    if (i == 3) x = (z[n - 1] - z[n]) * 1.6Q + z[n - 1];
    if (i > 3) x = (z[m + 1] - z[m + 2]) * C + z[m + 3];
    if (i == k && ind == 1) x = 0;
    int niter = 0;
    qf check = 1e-32Q;
    for (;;) {
      pb = 1;
      niter++;
      if (niter > 100) check = check * 10;
      pc = x;
      qf dj = A;
      for (int j = 2; j <= n; j++) {
        dj = dj + A;
        pa = pb;
        pb = pc;
        pc = x * pb + (x * pb - pa) * (dj - A) / dj;
      }
      pa = A / ((pb - x * pc) * f);
      pb = pa * pc * (A - x * x);
      x = x - pb;
      if (!(fabsq(pb) > check * fabsq(x))) break;
    }
    z[m] = x;
    w[m] = pa * pa * (A - x * x);
    w[m] = B * w[m];
    if (i == k && ind == 1) continue;
    z[i] = -z[m];
    w[i] = w[m];
  }
}

/* CONST (tmatrix_amplq.lp.f:972-1023), NP = -1.  sqrt(float(2n+1)/float(nn))
 * is a single-precision expression there. */
static void tm_const(tm_work *t, int ngauss, int nmax) {
  float dd[TM_NPN1 + 1];
  for (int n = 1; n <= nmax; n++) {
    int nn = n * (n + 1);
    t->an[n] = (qf)(float)nn;
    float d = sqrtf((float)(2 * n + 1) / (float)nn);
    dd[n] = d;
    for (int n1 = 1; n1 <= n; n1++) {
      qf ddd = (qf)d * (qf)dd[n1] * 0.5Q;
      t->ann[n][n1] = ddd;
      t->ann[n1][n] = ddd;
    }
  }
  int ng = 2 * ngauss;
  tm_qgauss(ng, t->x, t->w);
  for (int i = 1; i <= ngauss; i++) {
    qf y = t->x[i];
    y = 1 / (1 - y * y);
    t->ss[i] = y;
    t->ss[ng - i + 1] = y;
    y = sqrtq(y);
    t->s[i] = y;
    t->s[ng - i + 1] = y;
  }
}

/* RSP1 (tmatrix_amplq.lp.f:1194-1225): r(y)^2 and r'(y)/r(y) of a spheroid
 * of equal-volume radius rev and axis ratio eps. */
static void tm_rsp1(tm_work *t, int ng, int ngauss, qf rev, qf eps) {
  qf a = rev * powq(eps, 1 / (qf)3);
  qf aa = a * a, ee = eps * eps, ee1 = ee - 1;
  for (int i = 1; i <= ngauss; i++) {
    qf c = t->x[i], cc = c * c, ss = 1 - cc, s = sqrtq(ss);
    qf rr = 1 / (ss + ee * cc);
    t->r[i] = aa * rr;
    t->r[ng - i + 1] = t->r[i];
    t->dr[i] = rr * c * s * ee1;
    t->dr[ng - i + 1] = -t->dr[i];
  }
}

/* RJB (tmatrix_amplq.lp.f:1383-1410): j_n(x) and (1/x) d/dx (x j_n) by
 * downward recurrence of the ratios from order nmax + nnmax. */
static void tm_rjb(qf x, qf *y, qf *u, int nmax, int nnmax) {
  int l = nmax + nnmax;
  qf xx = 1 / x;
  qf *z = (qf *)malloc(sizeof(qf) * (l + 1));
  z[l] = 1 / ((qf)(2 * l + 1) * xx);
  for (int i = 1; i <= l - 1; i++) {
    int i1 = l - i;
    z[i1] = 1 / ((qf)(2 * i1 + 1) * xx - z[i1 + 1]);
  }
  qf z0 = 1 / (xx - z[1]);
  qf y0 = z0 * cosq(x) * xx;
  qf y1 = y0 * z[1];
  u[1] = y0 - y1 * xx;
  y[1] = y1;
  for (int i = 2; i <= nmax; i++) {
    qf yi1 = y[i - 1], yi = yi1 * z[i];
    u[i] = yi1 - (qf)i * yi * xx;
    y[i] = yi;
  }
  free(z);
}

/* RYB (tmatrix_amplq.lp.f:1416-1437): y_n(x) by upward recurrence. */
static void tm_ryb(qf x, qf *y, qf *v, int nmax) {
  qf c = cosq(x), s = sinq(x);
  qf x1 = 1 / x, x2 = x1 * x1, x3 = x2 * x1;
  qf y1 = -c * x2 - s * x1;
  y[1] = y1;
  y[2] = (-3 * x3 + x1) * c - 3 * x2 * s;
  for (int i = 2; i <= nmax - 1; i++) y[i + 1] = (qf)(2 * i + 1) * x1 * y[i] - y[i - 1];
  v[1] = -x1 * (c + y1);
  for (int i = 2; i <= nmax; i++) v[i] = y[i - 1] - (qf)i * x1 * y[i];
}

/* CJB (tmatrix_amplq.lp.f:1445-1520): j_n of the complex argument xr + i xi. */
static void tm_cjb(qf xr, qf xi, qf *yr, qf *yi, qf *ur, qf *ui, int nmax, int nnmax) {
  int l = nmax + nnmax;
  qf *czr = (qf *)malloc(sizeof(qf) * 2 * (l + 1)), *czi = czr + (l + 1);
  qf xrxi = 1 / (xr * xr + xi * xi);
  qf cxxr = xr * xrxi, cxxi = -xi * xrxi;
  qf qf_ = 1 / (qf)(2 * l + 1);
  czr[l] = xr * qf_;
  czi[l] = xi * qf_;
  for (int i = 1; i <= l - 1; i++) {
    int i1 = l - i;
    qf_ = (qf)(2 * i1 + 1);
    qf ar = qf_ * cxxr - czr[i1 + 1], ai = qf_ * cxxi - czi[i1 + 1];
    qf ari = 1 / (ar * ar + ai * ai);
    czr[i1] = ar * ari;
    czi[i1] = -ai * ari;
  }
  qf ar = cxxr - czr[1], ai = cxxi - czi[1];
  qf ari = 1 / (ar * ar + ai * ai);
  qf cz0r = ar * ari, cz0i = -ai * ari;
  qf cr = cosq(xr) * coshq(xi), ci = -sinq(xr) * sinhq(xi);
  ar = cz0r * cr - cz0i * ci;
  ai = cz0i * cr + cz0r * ci;
  qf cy0r = ar * cxxr - ai * cxxi, cy0i = ai * cxxr + ar * cxxi;
  qf cy1r = cy0r * czr[1] - cy0i * czi[1], cy1i = cy0i * czr[1] + cy0r * czi[1];
  ar = cy1r * cxxr - cy1i * cxxi;
  ai = cy1i * cxxr + cy1r * cxxi;
  yr[1] = cy1r;
  yi[1] = cy1i;
  ur[1] = cy0r - ar;
  ui[1] = cy0i - ai;
  for (int i = 2; i <= nmax; i++) {
    qf q = i;
    qf c1r = yr[i - 1], c1i = yi[i - 1];
    qf cr_ = c1r * czr[i] - c1i * czi[i], ci_ = c1i * czr[i] + c1r * czi[i];
    ar = cr_ * cxxr - ci_ * cxxi;
    ai = ci_ * cxxr + cr_ * cxxi;
    yr[i] = cr_;
    yi[i] = ci_;
    ur[i] = c1r - q * ar;
    ui[i] = c1i - q * ai;
  }
  free(czr);
}

/* VARY (tmatrix_amplq.lp.f:1114-1180), NP = -1.  The Bessel functions are
 * needed at the first NGAUSS points only (NCHECK = 1: TMATR0 and TMATR sum
 * over those); the other half repeats them.  Returns 0 or PAM_FATAL. */
static int tm_vary(tm_work *t, qf lam, qf mrr, qf mri, qf a, qf eps, int ngauss, qf p, int nmax) {
  int ng = ngauss * 2;
  tm_rsp1(t, ng, ngauss, a, eps);
  qf pi = p * 2 / lam;
  t->ppi = pi * pi;
  t->pir = t->ppi * mrr;
  t->pii = t->ppi * mri;
  qf v = 1 / (mrr * mrr + mri * mri);
  qf prr = mrr * v, pri = -mri * v;
  qf ta = 0;
  qf z[TM_NPNG2 + 1], zr[TM_NPNG2 + 1], zi[TM_NPNG2 + 1];
  for (int i = 1; i <= ng; i++) {
    qf vv = sqrtq(t->r[i]);
    v = vv * pi;
    if (v > ta) ta = v;
    vv = 1 / v;
    t->ddr[i] = vv;
    t->drr[i] = prr * vv;
    t->dri[i] = pri * vv;
    z[i] = v;
    zr[i] = v * mrr;
    zi[i] = v * mri;
  }
  qf tb = ta * sqrtq(mrr * mrr + mri * mri);
  qf fn = (qf)(float)nmax;
  if (fn > tb) tb = fn;
  int nnmax1 = (int)(8.0Q * sqrtq(ta > fn ? ta : fn) + 3);
  int nnmax2 = (int)(tb + 4 * powq(tb, 0.33333Q) + 8.0Q * sqrtq(tb));
  nnmax2 = nnmax2 - nmax + 5;
  if (nmax + nnmax2 >= 1200) return PAM_FATAL;
  for (int i = 1; i <= ngauss; i++) {
    tm_rjb(z[i], t->bj[i], t->bdj[i], nmax, nnmax1);
    tm_ryb(z[i], t->by[i], t->bdy[i], nmax);
    tm_cjb(zr[i], zi[i], t->bjr[i], t->bji[i], t->bdjr[i], t->bdji[i], nmax, nnmax2);
  }
  return 0;
}

/* VIG (tmatrix_amplq.lp.f:2049-2097): d^n_{0m}(arccos x) and its derivative
 * for n = 1..nmax.  sqrt(float(i2-1)/float(i2)) is single precision there. */
static void tm_vig(qf x, int nmax, int m, qf *dv1, qf *dv2) {
  qf a = 1;
  qf qs = sqrtq(1 - x * x), qs1 = 1 / qs;
  for (int n = 1; n <= nmax; n++) dv1[n] = dv2[n] = 0;
  if (m == 0) {
    qf d1 = 1, d2 = x;
    for (int n = 1; n <= nmax; n++) {
      qf qn = n, qn1 = n + 1, qn2 = 2 * n + 1;
      qf d3 = (qn2 * x * d2 - qn * d1) / qn1;
      qf der = qs1 * (qn1 * qn / qn2) * (-d1 + d3);
      dv1[n] = d2;
      dv2[n] = der;
      d1 = d2;
      d2 = d3;
    }
    return;
  }
  qf qmm = (qf)(m * m);
  for (int i = 1; i <= m; i++) {
    int i2 = i * 2;
    a = a * (qf)sqrtf((float)(i2 - 1) / (float)i2) * qs;
  }
  qf d1 = 0, d2 = a;
  for (int n = m; n <= nmax; n++) {
    qf qn = n, qn2 = 2 * n + 1, qn1 = n + 1;
    qf qnm = sqrtq(qn * qn - qmm), qnm1 = sqrtq(qn1 * qn1 - qmm);
    qf d3 = (qn2 * x * d2 - qnm * d1) / qnm1;
    qf der = qs1 * (-(qn1 * qnm * d1) + qn * qnm1 * d3) / qn2;
    dv1[n] = d2;
    dv2[n] = der;
    d1 = d2;
    d2 = d3;
  }
}

/* In-place inverse of the complex n x n matrix (ar, ai) by LU factorisation
 * with partial pivoting (the ZGETRF/ZGETRI pair of tmatrix_lpq.f, pivot by
 * |re| + |im| as IZAMAX), binary128.  Returns nonzero when singular. */
static int tm_cinv(int n, qf ar[][TM_NPN2], qf ai[][TM_NPN2]) {
  int piv[TM_NPN2];
  for (int k = 0; k < n; k++) {
    int p = k;
    qf best = -1;
    for (int i = k; i < n; i++) {
      qf v = fabsq(ar[i][k]) + fabsq(ai[i][k]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
    piv[k] = p;
    if (best == 0) return 1;
    if (p != k)
      for (int j = 0; j < n; j++) {
        qf tr = ar[k][j], ti = ai[k][j];
        ar[k][j] = ar[p][j];
        ai[k][j] = ai[p][j];
        ar[p][j] = tr;
        ai[p][j] = ti;
      }
    /* column below the pivot times 1/pivot, then the rank-1 update */
    qf den = ar[k][k] * ar[k][k] + ai[k][k] * ai[k][k];
    qf pr = ar[k][k] / den, pi = -ai[k][k] / den;
    for (int i = k + 1; i < n; i++) {
      qf xr = ar[i][k] * pr - ai[i][k] * pi, xi = ar[i][k] * pi + ai[i][k] * pr;
      ar[i][k] = xr;
      ai[i][k] = xi;
      for (int j = k + 1; j < n; j++) {
        ar[i][j] -= xr * ar[k][j] - xi * ai[k][j];
        ai[i][j] -= xr * ai[k][j] + xi * ar[k][j];
      }
    }
  }
  /* inverse of U in place */
  for (int j = 0; j < n; j++) {
    qf den = ar[j][j] * ar[j][j] + ai[j][j] * ai[j][j];
    qf dr = ar[j][j] / den, di = -ai[j][j] / den;
    ar[j][j] = dr;
    ai[j][j] = di;
    qf ajr = -dr, aji = -di;
    for (int i = 0; i < j; i++) {
      qf sr = 0, si = 0;
      for (int k = i; k < j; k++) {
        sr += ar[i][k] * ar[k][j] - ai[i][k] * ai[k][j];
        si += ar[i][k] * ai[k][j] + ai[i][k] * ar[k][j];
      }
      ar[i][j] = sr;
      ai[i][j] = si;
    }
    for (int i = 0; i < j; i++) {
      qf xr = ar[i][j] * ajr - ai[i][j] * aji, xi = ar[i][j] * aji + ai[i][j] * ajr;
      ar[i][j] = xr;
      ai[i][j] = xi;
    }
  }
  /* inv(A) = inv(U) inv(L), columns right to left, then the column swaps */
  qf wr[TM_NPN2], wi[TM_NPN2];
  for (int j = n - 2; j >= 0; j--) {
    for (int i = j + 1; i < n; i++) {
      wr[i] = ar[i][j];
      wi[i] = ai[i][j];
      ar[i][j] = ai[i][j] = 0;
    }
    for (int i = 0; i < n; i++) {
      qf sr = 0, si = 0;
      for (int k = j + 1; k < n; k++) {
        sr += ar[i][k] * wr[k] - ai[i][k] * wi[k];
        si += ar[i][k] * wi[k] + ai[i][k] * wr[k];
      }
      ar[i][j] -= sr;
      ai[i][j] -= si;
    }
  }
  for (int j = n - 2; j >= 0; j--) {
    int p = piv[j];
    if (p != j)
      for (int i = 0; i < n; i++) {
        qf tr = ar[i][j], ti = ai[i][j];
        ar[i][j] = ar[i][p];
        ai[i][j] = ai[i][p];
        ar[i][p] = tr;
        ai[i][p] = ti;
      }
  }
  return 0;
}

/* TT (tmatrix_amplq.lp.f:2107-2150): T = -RG(Q) Q^-1 for the 2 nm square
 * matrices.  ZQ = cmplx(QR, QI) has no KIND there, so Q enters the
 * binary128 inversion rounded to single precision (default complex); the
 * product is double (REAL*8 implicit typing of TT: each operand rounded to
 * double, summed over k in order). */
static int tm_tt(tm_work *t, int nm) {
  int n = 2 * nm;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      t->qr[i][j] = (qf)(float)t->qr[i][j];
      t->qi[i][j] = (qf)(float)t->qi[i][j];
    }
  if (tm_cinv(n, t->qr, t->qi)) return PAM_FATAL;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double tr = 0.0, ti = 0.0;
      for (int k = 0; k < n; k++) {
        double arr = (double)t->rgqr[i][k], ari = (double)t->rgqi[i][k];
        double ar = (double)t->qr[k][j], ai = (double)t->qi[k][j];
        tr = tr - arr * ar + ari * ai;
        ti = ti - arr * ai - ari * ar;
      }
      t->tr1[i][j] = tr;
      t->ti1[i][j] = ti;
    }
  return 0;
}

/* TMATR0 and TMATR (tmatrix_amplq.lp.f:1526-2040) with NCHECK = 1: the Q and
 * RG(Q) matrices of azimuthal order m from the Gauss sums over the first
 * NGAUSS points, then TT.  m = 0 has only the 12/21 sums and zero
 * off-diagonal blocks. */
static int tm_tmatr(tm_work *t, int m, int ngauss, int nmax) {
  const int ngss = ngauss;
  const qf factor = 2;
  qf sig[2 * TM_NPN1 + 2];
  qf si_ = 1;
  for (int n = 1; n <= 2 * nmax; n++) {
    si_ = -si_;
    sig[n] = si_;
  }
  qf dv1[TM_NPN1 + 1], dv2[TM_NPN1 + 1];
  for (int i = 1; i <= ngauss; i++) {
    int i1 = ngauss + i, i2 = ngauss - i + 1;
    tm_vig(t->x[i1], nmax, m, dv1, dv2);
    for (int n = 1; n <= nmax; n++) {
      qf si = sig[n + m];
      t->d1[i1][n] = dv1[n];
      t->d2[i1][n] = dv2[n];
      t->d1[i2][n] = dv1[n] * si;
      t->d2[i2][n] = -dv2[n] * si;
    }
  }
  qf rr[TM_NPNG1 + 1], ds[TM_NPNG1 + 1], dss[TM_NPNG1 + 1];
  const qf qm = m, qmm = qm * qm;
  for (int i = 1; i <= ngss; i++) {
    qf wr = t->w[i] * t->r[i];
    ds[i] = t->s[i] * qm * wr;
    dss[i] = t->ss[i] * qmm;
    rr[i] = wr;
  }
  const int mm1 = m > 0 ? m : 1;
  const int nm = nmax - mm1 + 1;
  for (int n1 = mm1; n1 <= nmax; n1++) {
    qf an1 = t->an[n1];
    for (int n2 = mm1; n2 <= nmax; n2++) {
      qf an2 = t->an[n2];
      qf ar11 = 0, ar12 = 0, ar21 = 0, ar22 = 0, ai11 = 0, ai12 = 0, ai21 = 0, ai22 = 0;
      qf gr11 = 0, gr12 = 0, gr21 = 0, gr22 = 0, gi11 = 0, gi12 = 0, gi21 = 0, gi22 = 0;
      qf si = sig[n1 + n2];
      const int even = si > 0; /* n1 + n2 even: 12 and 21 sums, odd: 11 and 22 */
      if (m == 0 && !even) goto store;
      for (int i = 1; i <= ngss; i++) {
        qf d1n1 = t->d1[i][n1], d2n1 = t->d2[i][n1], d1n2 = t->d1[i][n2], d2n2 = t->d2[i][n2];
        qf a11 = d1n1 * d1n2, a12 = d1n1 * d2n2, a21 = d2n1 * d1n2, a22 = d2n1 * d2n2;
        qf aa1 = a12 + a21;
        qf qj1 = t->bj[i][n1], qy1 = t->by[i][n1];
        qf qjr2 = t->bjr[i][n2], qji2 = t->bji[i][n2];
        qf qdjr2 = t->bdjr[i][n2], qdji2 = t->bdji[i][n2];
        qf qdj1 = t->bdj[i][n1], qdy1 = t->bdy[i][n1];
        qf c1r = qjr2 * qj1, c1i = qji2 * qj1;
        qf b1r = c1r - qji2 * qy1, b1i = c1i + qjr2 * qy1;
        qf c2r = qjr2 * qdj1, c2i = qji2 * qdj1;
        qf b2r = c2r - qji2 * qdy1, b2i = c2i + qjr2 * qdy1;
        qf ddri = t->ddr[i];
        qf c3r = ddri * c1r, c3i = ddri * c1i, b3r = ddri * b1r, b3i = ddri * b1i;
        qf c4r = qdjr2 * qj1, c4i = qdji2 * qj1;
        qf b4r = c4r - qdji2 * qy1, b4i = c4i + qdjr2 * qy1;
        qf drri = t->drr[i], drii = t->dri[i];
        qf c5r = c1r * drri - c1i * drii, c5i = c1i * drri + c1r * drii;
        qf b5r = b1r * drri - b1i * drii, b5i = b1i * drri + b1r * drii;
        qf uri = t->dr[i], rri = rr[i];
        if (m == 0) {
          qf f1 = rri * a22;
          qf f2 = rri * uri * an1 * a12;
          ar12 = ar12 + f1 * b2r + f2 * b3r;
          ai12 = ai12 + f1 * b2i + f2 * b3i;
          gr12 = gr12 + f1 * c2r + f2 * c3r;
          gi12 = gi12 + f1 * c2i + f2 * c3i;
          f2 = rri * uri * an2 * a21;
          ar21 = ar21 + f1 * b4r + f2 * b5r;
          ai21 = ai21 + f1 * b4i + f2 * b5i;
          gr21 = gr21 + f1 * c4r + f2 * c5r;
          gi21 = gi21 + f1 * c4i + f2 * c5i;
          continue;
        }
        qf aa2 = a11 * dss[i] + a22;
        qf dsi = ds[i];
        if (even) {
          qf f1 = rri * aa2;
          qf f2 = rri * uri * an1 * a12;
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
          qf c6r = qdjr2 * qdj1, c6i = qdji2 * qdj1;
          qf b6r = c6r - qdji2 * qdy1, b6i = c6i + qdjr2 * qdy1;
          qf c7r = c4r * ddri, c7i = c4i * ddri, b7r = b4r * ddri, b7i = b4i * ddri;
          qf c8r = c2r * drri - c2i * drii, c8i = c2i * drri + c2r * drii;
          qf b8r = b2r * drri - b2i * drii, b8i = b2i * drri + b2r * drii;
          qf e1 = dsi * aa1;
          ar11 = ar11 + e1 * b1r;
          ai11 = ai11 + e1 * b1i;
          gr11 = gr11 + e1 * c1r;
          gi11 = gi11 + e1 * c1i;
          qf e2 = dsi * uri * a11;
          qf e3 = e2 * an2;
          e2 = e2 * an1;
          ar22 = ar22 + e1 * b6r + e2 * b7r + e3 * b8r;
          ai22 = ai22 + e1 * b6i + e2 * b7i + e3 * b8i;
          gr22 = gr22 + e1 * c6r + e2 * c7r + e3 * c8r;
          gi22 = gi22 + e1 * c6i + e2 * c7i + e3 * c8i;
        }
      }
    store:;
      qf an12 = t->ann[n1][n2] * factor;
      qf r11 = ar11 * an12, r12 = ar12 * an12, r21 = ar21 * an12, r22 = ar22 * an12;
      qf i11 = ai11 * an12, i12 = ai12 * an12, i21 = ai21 * an12, i22 = ai22 * an12;
      qf rg11 = gr11 * an12, rg12 = gr12 * an12, rg21 = gr21 * an12, rg22 = gr22 * an12;
      qf ig11 = gi11 * an12, ig12 = gi12 * an12, ig21 = gi21 * an12, ig22 = gi22 * an12;
      /* Q entries (the 310 loops) */
      const qf tpir = t->pir, tpii = t->pii, tppi = t->ppi;
      int k1 = n1 - mm1, kk1 = k1 + nm, k2 = n2 - mm1, kk2 = k2 + nm;
      qf tar11 = -r11, tai11 = -i11, tgr11 = -rg11, tgi11 = -ig11;
      qf tar12 = i12, tai12 = -r12, tgr12 = ig12, tgi12 = -rg12;
      qf tar21 = -i21, tai21 = r21, tgr21 = -ig21, tgi21 = rg21;
      qf tar22 = -r22, tai22 = -i22, tgr22 = -rg22, tgi22 = -ig22;
      t->qr[k1][k2] = tpir * tar21 - tpii * tai21 + tppi * tar12;
      t->qi[k1][k2] = tpir * tai21 + tpii * tar21 + tppi * tai12;
      t->rgqr[k1][k2] = tpir * tgr21 - tpii * tgi21 + tppi * tgr12;
      t->rgqi[k1][k2] = tpir * tgi21 + tpii * tgr21 + tppi * tgi12;
      t->qr[kk1][kk2] = tpir * tar12 - tpii * tai12 + tppi * tar21;
      t->qi[kk1][kk2] = tpir * tai12 + tpii * tar12 + tppi * tai21;
      t->rgqr[kk1][kk2] = tpir * tgr12 - tpii * tgi12 + tppi * tgr21;
      t->rgqi[kk1][kk2] = tpir * tgi12 + tpii * tgr12 + tppi * tgi21;
      if (m == 0) {
        t->qr[k1][kk2] = t->qi[k1][kk2] = t->rgqr[k1][kk2] = t->rgqi[k1][kk2] = 0;
        t->qr[kk1][k2] = t->qi[kk1][k2] = t->rgqr[kk1][k2] = t->rgqi[kk1][k2] = 0;
      } else {
        t->qr[k1][kk2] = tpir * tar11 - tpii * tai11 + tppi * tar22;
        t->qi[k1][kk2] = tpir * tai11 + tpii * tar11 + tppi * tai22;
        t->rgqr[k1][kk2] = tpir * tgr11 - tpii * tgi11 + tppi * tgr22;
        t->rgqi[k1][kk2] = tpir * tgi11 + tpii * tgr11 + tppi * tgi22;
        t->qr[kk1][k2] = tpir * tar22 - tpii * tai22 + tppi * tar11;
        t->qi[kk1][k2] = tpir * tai22 + tpii * tar22 + tppi * tai11;
        t->rgqr[kk1][k2] = tpir * tgr22 - tpii * tgi22 + tppi * tgr11;
        t->rgqi[kk1][k2] = tpir * tgi22 + tpii * tgr22 + tppi * tgi11;
      }
    }
  }
  return tm_tt(t, nm);
}

/* QEXT and QSCA of the m = 0 matrix (the convergence test of
 * tmatrix_amplq.lp.f:388-405), double */
static void tm_q0(const tm_work *t, int nmax, double *qext, double *qsca) {
  double qe = 0.0, qs = 0.0;
  for (int n = 1; n <= nmax; n++) {
    int n1 = n + nmax;
    double tr = t->tr1[n - 1][n - 1], ti = t->ti1[n - 1][n - 1];
    double tr1 = t->tr1[n1 - 1][n1 - 1], ti1 = t->ti1[n1 - 1][n1 - 1];
    double dn1 = (double)(2 * n + 1);
    qs = qs + dn1 * (tr * tr + ti * ti + tr1 * tr1 + ti1 * ti1);
    qe = qe + (tr + tr1) * dn1;
  }
  *qext = qe;
  *qsca = qs;
}

/* tmatrix_amplq (tmatrix_amplq.lp.f:276-540): the T matrix of a spheroid of
 * equal-volume radius axi [m] and axis ratio as_ratio (EPS = 1/as_ratio) at
 * wavelength lam, refractive index mrr + i mri.  NMAX grows from
 * max(4, x + 4.05 x^(1/3)) until Qext and Qsca of T(0) change by at most
 * 1e-3, then NGAUSS from 2 NMAX + 1 likewise.  Fills T (allocated here) and
 * returns 0 or PAM_FATAL. */
static int tm_solve(tm_work *t, double lam_d, double mrr_d, double mri_d, double axi,
                    double as_ratio, tm_tmat *T) {
  const qf lam = lam_d, mrr = mrr_d, mri = mri_d;
  const qf eps = (qf)(1 / as_ratio);
  const double ddelt = 0.001;
  const int ndgs = 2;
  const qf p = acosq(-1);
  const double pin = (double)p;
  const double dlam = (double)lam;
  const qf a = (qf)(1.0 * axi);
  const double xev = (double)((qf)(2.0 * pin) * a / (qf)dlam);
  int ixxx = (int)(xev + 4.05 * pow(xev, 0.333333));
  int inm1 = ixxx > 4 ? ixxx : 4;
  if (inm1 >= TM_NPN1) return PAM_FATAL;
  double qext1 = 0.0, qsca1 = 0.0;
  int nmax = 0, ngauss = 0, conv = 0;
  for (int nma = inm1; nma <= TM_NPN1; nma++) {
    nmax = nma;
    ngauss = nmax * ndgs;
    if (ngauss > TM_NPNG1) return PAM_FATAL;
    tm_const(t, ngauss, nmax);
    if (tm_vary(t, lam, mrr, mri, a, eps, ngauss, p, nmax)) return PAM_FATAL;
    if (tm_tmatr(t, 0, ngauss, nmax)) return PAM_FATAL;
    double qext, qsca;
    tm_q0(t, nmax, &qext, &qsca);
    double dsca = fabs((qsca1 - qsca) / qsca), dext = fabs((qext1 - qext) / qext);
    qext1 = qext;
    qsca1 = qsca;
    if (dsca <= ddelt && dext <= ddelt) {
      conv = 1;
      break;
    }
  }
  if (!conv) return PAM_FATAL;
  if (ngauss != TM_NPNG1) {
    for (int ngaus = ngauss + 1; ngaus <= TM_NPNG1; ngaus++) {
      ngauss = ngaus;
      tm_const(t, ngauss, nmax);
      if (tm_vary(t, lam, mrr, mri, a, eps, ngauss, p, nmax)) return PAM_FATAL;
      if (tm_tmatr(t, 0, ngauss, nmax)) return PAM_FATAL;
      double qext, qsca;
      tm_q0(t, nmax, &qext, &qsca);
      double dsca = fabs((qsca1 - qsca) / qsca), dext = fabs((qext1 - qext) / qext);
      qext1 = qext;
      qsca1 = qsca;
      if (dsca <= ddelt && dext <= ddelt) break;
    }
  }
  T->nmax = nmax;
  T->t = (float *)calloc((size_t)8 * (nmax + 1) * nmax * nmax, sizeof(float));
  for (int n2 = 1; n2 <= nmax; n2++)
    for (int n1 = 1; n1 <= nmax; n1++) {
      int nn1 = n1 + nmax, nn2 = n2 + nmax;
      *tm_at(T, 0, 0, n1, n2) = (float)t->tr1[n1 - 1][n2 - 1];
      *tm_at(T, 4, 0, n1, n2) = (float)t->ti1[n1 - 1][n2 - 1];
      *tm_at(T, 1, 0, n1, n2) = (float)t->tr1[n1 - 1][nn2 - 1];
      *tm_at(T, 5, 0, n1, n2) = (float)t->ti1[n1 - 1][nn2 - 1];
      *tm_at(T, 2, 0, n1, n2) = (float)t->tr1[nn1 - 1][n2 - 1];
      *tm_at(T, 6, 0, n1, n2) = (float)t->ti1[nn1 - 1][n2 - 1];
      *tm_at(T, 3, 0, n1, n2) = (float)t->tr1[nn1 - 1][nn2 - 1];
      *tm_at(T, 7, 0, n1, n2) = (float)t->ti1[nn1 - 1][nn2 - 1];
    }
  for (int m = 1; m <= nmax; m++) {
    if (tm_tmatr(t, m, ngauss, nmax)) {
      free(T->t);
      T->t = NULL;
      return PAM_FATAL;
    }
    int nm = nmax - m + 1;
    for (int n2 = 1; n2 <= nm; n2++)
      for (int n1 = 1; n1 <= nm; n1++) {
        int nn1 = n1 + m - 1, nn2 = n2 + m - 1, n11 = n1 + nm, n22 = n2 + nm;
        *tm_at(T, 0, m, nn1, nn2) = (float)t->tr1[n1 - 1][n2 - 1];
        *tm_at(T, 4, m, nn1, nn2) = (float)t->ti1[n1 - 1][n2 - 1];
        *tm_at(T, 1, m, nn1, nn2) = (float)t->tr1[n1 - 1][n22 - 1];
        *tm_at(T, 5, m, nn1, nn2) = (float)t->ti1[n1 - 1][n22 - 1];
        *tm_at(T, 2, m, nn1, nn2) = (float)t->tr1[n11 - 1][n2 - 1];
        *tm_at(T, 6, m, nn1, nn2) = (float)t->ti1[n11 - 1][n2 - 1];
        *tm_at(T, 3, m, nn1, nn2) = (float)t->tr1[n11 - 1][n22 - 1];
        *tm_at(T, 7, m, nn1, nn2) = (float)t->ti1[n11 - 1][n22 - 1];
      }
  }
  return 0;
}

/* VIGAMPL (tmatrix_amplq.lp.f:905-960): d^n_{0m}(arccos x)/sin and the
 * derivative, double; at |x| = 1 only m = 1 is nonzero. */
static void tm_vigampl(double x, int nmax, int m, double *dv1, double *dv2) {
  for (int n = 1; n <= nmax; n++) dv1[n] = dv2[n] = 0.0;
  double dx = fabs(x);
  if (fabs(1.0 - dx) <= 1e-10) {
    if (m != 1) return;
    for (int n = 1; n <= nmax; n++) {
      double dn = (double)(n * (n + 1));
      dn = 0.5 * sqrt(dn);
      if (x < 0.0) dn = dn * ((n + 1) % 2 == 0 ? 1.0 : -1.0);
      dv1[n] = dn;
      if (x < 0.0) dn = -dn;
      dv2[n] = dn;
    }
    return;
  }
  double a = 1.0, qs = sqrt(1.0 - x * x), qs1 = 1.0 / qs, dsi = qs1;
  if (m == 0) {
    double d1 = 1.0, d2 = x;
    for (int n = 1; n <= nmax; n++) {
      double qn = n, qn1 = n + 1, qn2 = 2 * n + 1;
      double d3 = (qn2 * x * d2 - qn * d1) / qn1;
      double der = qs1 * (qn1 * qn / qn2) * (-d1 + d3);
      dv1[n] = d2 * dsi;
      dv2[n] = der;
      d1 = d2;
      d2 = d3;
    }
    return;
  }
  double qmm = (double)(m * m);
  for (int i = 1; i <= m; i++) {
    int i2 = i * 2;
    a = a * sqrt((double)(i2 - 1) / (double)i2) * qs;
  }
  double d1 = 0.0, d2 = a;
  for (int n = m; n <= nmax; n++) {
    double qn = n, qn2 = 2 * n + 1, qn1 = n + 1;
    double qnm = sqrt(qn * qn - qmm), qnm1 = sqrt(qn1 * qn1 - qmm);
    double d3 = (qn2 * x * d2 - qnm * d1) / qnm1;
    double der = qs1 * (-(qn1 * qnm * d1) + qn * qnm1 * d3) / qn2;
    dv1[n] = d2 * dsi;
    dv2[n] = der;
    d1 = d2;
    d2 = d3;
  }
}

typedef struct {
  double re, im;
} zc;
static inline zc zmul(zc a, zc b) { return (zc){a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
static inline zc zadd(zc a, zc b) { return (zc){a.re + b.re, a.im + b.im}; }
static inline zc zsub(zc a, zc b) { return (zc){a.re - b.re, a.im - b.im}; }
static inline zc zscl(zc a, double s) { return (zc){a.re * s, a.im * s}; }

/* tmatrix_AMPL (tmatrix_amplq.lp.f:602-888): amplitude matrix (VV, VH, HV,
 * HH) for incidence (tl, pl) and scattering (tl1, pl1) in degrees, particle
 * orientation (alpha, beta), from the stored REAL*4 T matrix.  out[8] = re,
 * im of VV, VH, HV, HH. */
static void tm_ampl(const tm_tmat *T, double dlam, double tl, double tl1, double pl, double pl1,
                    double alpha, double beta, double out[8]) {
  const int nmax = T->nmax;
  const double pin = acos(-1.0), pin2 = pin * 0.5, pi = pin / 180.0;
  double alph = alpha * pi, bet = beta * pi;
  double thetl = tl * pi, phil = pl * pi, thetl1 = tl1 * pi, phil1 = pl1 * pi;
  const double eps = 1e-8;
  if (thetl < pin2) thetl = thetl + eps;
  if (thetl > pin2) thetl = thetl - eps;
  if (thetl1 < pin2) thetl1 = thetl1 + eps;
  if (thetl1 > pin2) thetl1 = thetl1 - eps;
  if (phil < pin) phil = phil + eps;
  if (phil > pin) phil = phil - eps;
  if (phil1 < pin) phil1 = phil1 + eps;
  if (phil1 > pin) phil1 = phil1 - eps;
  if (bet <= pin2 && pin2 - bet <= eps) bet = bet - eps;
  if (bet > pin2 && bet - pin2 <= eps) bet = bet + eps;
  double cb = cos(bet), sb = sin(bet), ct = cos(thetl), st = sin(thetl);
  double cp = cos(phil - alph), sp = sin(phil - alph);
  double ctp = ct * cb + st * sb * cp;
  double thetp = acos(ctp);
  double cpp = cb * st * cp - sb * ct, spp = st * sp;
  double phip = atan(spp / cpp);
  if (phip > 0.0 && sp < 0.0) phip = phip + pin;
  if (phip < 0.0 && sp > 0.0) phip = phip + pin;
  if (phip < 0.0) phip = phip + 2.0 * pin;
  double ct1 = cos(thetl1), st1 = sin(thetl1);
  double cp1 = cos(phil1 - alph), sp1 = sin(phil1 - alph);
  double ctp1 = ct1 * cb + st1 * sb * cp1;
  double thetp1 = acos(ctp1);
  double cpp1 = cb * st1 * cp1 - sb * ct1, spp1 = st1 * sp1;
  double phip1 = atan(spp1 / cpp1);
  if (phip1 > 0.0 && sp1 < 0.0) phip1 = phip1 + pin;
  if (phip1 < 0.0 && sp1 > 0.0) phip1 = phip1 + pin;
  if (phip1 < 0.0) phip1 = phip1 + 2.0 * pin;
  double ca = cos(alph), sa = sin(alph);
  double B[3][3] = {{ca * cb, sa * cb, -sb}, {-sa, ca, 0.0}, {ca * sb, sa * sb, cb}};
  cp = cos(phil);
  sp = sin(phil);
  cp1 = cos(phil1);
  sp1 = sin(phil1);
  double AL[3][2] = {{ct * cp, -sp}, {ct * sp, cp}, {-st, 0.0}};
  double AL1[3][2] = {{ct1 * cp1, -sp1}, {ct1 * sp1, cp1}, {-st1, 0.0}};
  ct = ctp;
  st = sin(thetp);
  cp = cos(phip);
  sp = sin(phip);
  ct1 = ctp1;
  st1 = sin(thetp1);
  cp1 = cos(phip1);
  sp1 = sin(phip1);
  double AP[2][3] = {{ct * cp, ct * sp, -st}, {-sp, cp, 0.0}};
  double AP1[2][3] = {{ct1 * cp1, ct1 * sp1, -st1}, {-sp1, cp1, 0.0}};
  double C[3][2], R[2][2], R1[2][2];
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
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 2; j++) {
      double x = 0.0;
      for (int k = 0; k < 3; k++) x = x + B[i][k] * AL1[k][j];
      C[i][j] = x;
    }
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      double x = 0.0;
      for (int k = 0; k < 3; k++) x = x + AP1[i][k] * C[k][j];
      R1[i][j] = x;
    }
  double d = 1.0 / (R1[0][0] * R1[1][1] - R1[0][1] * R1[1][0]);
  double x = R1[0][0];
  R1[0][0] = R1[1][1] * d;
  R1[0][1] = -R1[0][1] * d;
  R1[1][0] = -R1[1][0] * d;
  R1[1][1] = x * d;
  /* CAL(n, nn) = i^(nn - n - 1) sqrt((2n+1)(2nn+1)/(n nn (n+1)(nn+1))) */
  zc *cal = (zc *)malloc(sizeof(zc) * (nmax + 1) * (nmax + 1));
  for (int nn = 1; nn <= nmax; nn++)
    for (int n = 1; n <= nmax; n++) {
      int e = ((nn - n - 1) % 4 + 4) % 4;
      zc cn = e == 0 ? (zc){1, 0} : e == 1 ? (zc){0, 1} : e == 2 ? (zc){-1, 0} : (zc){0, -1};
      double dnn = (double)((2 * n + 1) * (2 * nn + 1));
      dnn = dnn / (double)(n * nn * (n + 1) * (nn + 1));
      double rn = sqrt(dnn);
      cal[n * (nmax + 1) + nn] = zscl(cn, rn);
    }
  double dcth0 = ctp, dcth = ctp1, ph = phip1 - phip;
  zc vv = {0, 0}, vh = {0, 0}, hv = {0, 0}, hh = {0, 0};
  double *dv = (double *)malloc(sizeof(double) * 4 * (nmax + 1));
  double *dv1 = dv, *dv2 = dv + (nmax + 1), *dv01 = dv + 2 * (nmax + 1), *dv02 = dv + 3 * (nmax + 1);
  for (int m = 0; m <= nmax; m++) {
    int nmin = m > 1 ? m : 1;
    tm_vigampl(dcth, nmax, m, dv1, dv2);
    tm_vigampl(dcth0, nmax, m, dv01, dv02);
    double fc = 2.0 * cos(m * ph), fs = 2.0 * sin(m * ph);
    for (int nn = nmin; nn <= nmax; nn++) {
      double dv1nn = m * dv01[nn], dv2nn = dv02[nn];
      for (int n = nmin; n <= nmax; n++) {
        double dv1n = m * dv1[n], dv2n = dv2[n];
        zc ct11 = {*tm_at(T, 0, m, n, nn), *tm_at(T, 4, m, n, nn)};
        zc ct22 = {*tm_at(T, 3, m, n, nn), *tm_at(T, 7, m, n, nn)};
        zc c = cal[n * (nmax + 1) + nn];
        if (m == 0) {
          zc cn = zscl(zscl(c, dv2n), dv2nn);
          vv = zadd(vv, zmul(cn, ct22));
          hh = zadd(hh, zmul(cn, ct11));
        } else {
          zc ct12 = {*tm_at(T, 1, m, n, nn), *tm_at(T, 5, m, n, nn)};
          zc ct21 = {*tm_at(T, 2, m, n, nn), *tm_at(T, 6, m, n, nn)};
          zc cn1 = zscl(c, fc), cn2 = zscl(c, fs);
          double d11 = dv1n * dv1nn, d12 = dv1n * dv2nn, d21 = dv2n * dv1nn, d22 = dv2n * dv2nn;
          zc s;
          s = zadd(zadd(zadd(zscl(ct11, d11), zscl(ct21, d21)), zscl(ct12, d12)), zscl(ct22, d22));
          vv = zadd(vv, zmul(s, cn1));
          s = zadd(zadd(zadd(zscl(ct11, d12), zscl(ct21, d22)), zscl(ct12, d11)), zscl(ct22, d21));
          vh = zadd(vh, zmul(s, cn2));
          s = zadd(zadd(zadd(zscl(ct11, d21), zscl(ct21, d11)), zscl(ct12, d22)), zscl(ct22, d12));
          hv = zsub(hv, zmul(s, cn2));
          s = zadd(zadd(zadd(zscl(ct11, d22), zscl(ct21, d12)), zscl(ct12, d21)), zscl(ct22, d11));
          hh = zadd(hh, zmul(s, cn1));
        }
      }
    }
  }
  free(dv);
  free(cal);
  double dk = 2.0 * pin / dlam;
  vv = (zc){vv.re / dk, vv.im / dk};
  vh = (zc){vh.re / dk, vh.im / dk};
  hv = (zc){hv.re / dk, hv.im / dk};
  hh = (zc){hh.re / dk, hh.im / dk};
  zc cvv = zadd(zscl(vv, R[0][0]), zscl(vh, R[1][0]));
  zc cvh = zadd(zscl(vv, R[0][1]), zscl(vh, R[1][1]));
  zc chv = zadd(zscl(hv, R[0][0]), zscl(hh, R[1][0]));
  zc chh = zadd(zscl(hv, R[0][1]), zscl(hh, R[1][1]));
  vv = zadd(zscl(cvv, R1[0][0]), zscl(chv, R1[0][1]));
  vh = zadd(zscl(cvh, R1[0][0]), zscl(chh, R1[0][1]));
  hv = zadd(zscl(cvv, R1[1][0]), zscl(chv, R1[1][1]));
  hh = zadd(zscl(cvh, R1[1][0]), zscl(chh, R1[1][1]));
  out[0] = vv.re;
  out[1] = vv.im;
  out[2] = vh.re;
  out[3] = vh.im;
  out[4] = hv.re;
  out[5] = hv.im;
  out[6] = hh.re;
  out[7] = hh.im;
}

/* calc_single_tmatrix (tmatrix.f90:447-684), passive (qua_start = 1),
 * azimuth_num = 30, azimuth0_num = 1, alpha = 0.  m = mrr + i mri (the
 * conjugate the caller passes, imaginary part taken absolute).  Outputs in
 * the reference's Fortran order: smat(2,NMU,2,NMU,2), ext(2,2,NMU),
 * emis(2,NMU), sback(2,2) (left untouched when no angle is backward). */
static int tm_single(tm_work *t, double frequency, double mrr, double mri, double axi,
                     double as_ratio, double beta, double *smat, double *ext, double *emis,
                     double *sback) {
  const int azimuth_num = 30, azimuth0_num = 1;
  const double alpha = 0.0;
  double lam = C_LIGHT / frequency;
  double wave_num = 2.0 * PI_D / lam;
  if (!(as_ratio > 0.0) || !(axi > 0.0) || !(frequency > 0.0)) return PAM_FATAL;
  tm_tmat T = {0, NULL};
  if (tm_solve(t, lam, mrr, fabs(mri), axi, as_ratio, &T)) return PAM_FATAL;
  double mu[NMU], qw[NMU];
  pam_oracle_lobatto(NMU, mu, qw);
#define SM(a, l, b, j, k) smat[(a) + 2 * ((l) + NMU * ((b) + 2 * ((j) + NMU * (k))))]
#define EX(a, b, j) ext[(a) + 2 * ((b) + 2 * (j))]
  memset(smat, 0, sizeof(double) * 2 * NMU * 2 * NMU * 2);
  memset(ext, 0, sizeof(double) * 4 * NMU);
  memset(emis, 0, sizeof(double) * 2 * NMU);
  const double fact_sca = 0.5 / (wave_num * wave_num);
  const double fext = (2.0 * PI_D) / pow(wave_num, 2.0); /* fact_ext = (0, fext) */
  const float thet_max = 179.9999f;
  for (int jj = 0; jj < NMU; jj++) {
    double thet0 = acos(mu[jj] * 1.0) * 180.0 / PI_D;
    if (thet0 > (double)thet_max) thet0 = 180.0;
    double emis11[2 * NMU], emis12[2 * NMU];
    for (int kk = 0; kk < 2; kk++) {
      for (int ll = 0; ll < NMU; ll++) {
        double thet = acos(mu[ll] * (kk == 0 ? 1.0 : -1.0)) * 180.0 / PI_D;
        double thet_w = qw[ll];
        if (thet > (double)thet_max) thet = 180.0;
        for (int m = 0; m < azimuth0_num; m++) {
          double phi0 = 360.0 / (double)(float)azimuth0_num * ((double)(float)(m + 1) - 1.0);
          double phi0_w = 1.0 / 360.0 * (360.0 / azimuth0_num);
          double t11 = 0.0, t12 = 0.0, t21 = 0.0, t22 = 0.0;
          for (int n = 0; n < azimuth_num; n++) {
            double phi = 360.0 / (double)(float)azimuth_num * ((double)(float)(n + 1) - 1.0);
            double phi_w = 1.0 / 360.0 * (360.0 / azimuth_num);
            double s[8];
            tm_ampl(&T, lam, thet0, thet, phi0, phi, alpha, beta, s);
            for (int q = 0; q < 8; q++) s[q] = s[q] * wave_num;
            /* |s|^2 as the real part of s conj(s) */
            double a11 = s[0] * s[0] - s[1] * (-s[1]), a12 = s[2] * s[2] - s[3] * (-s[3]);
            double a21 = s[4] * s[4] - s[5] * (-s[5]), a22 = s[6] * s[6] - s[7] * (-s[7]);
            if (thet == 180.0 && phi == phi0 && sback) {
              /* Sback(1,1), (2,1), (1,2), (2,2) in Fortran order */
              sback[0] = fact_sca * (((a11 + a12) + a21) + a22);
              sback[1] = fact_sca * (((a11 + a12) - a21) - a22);
              sback[2] = fact_sca * (((a11 - a12) + a21) - a22);
              sback[3] = fact_sca * (((a11 - a12) - a21) + a22);
            }
            t11 = t11 + (fact_sca * (((a11 + a12) + a21) + a22)) * phi_w;
            t12 = t12 + (fact_sca * (((a11 - a12) + a21) - a22)) * phi_w;
            t21 = t21 + (fact_sca * (((a11 + a12) - a21) - a22)) * phi_w;
            t22 = t22 + (fact_sca * (((a11 - a12) - a21) + a22)) * phi_w;
            if (phi0 == phi && thet0 == thet) {
              /* -real((s11 +- s22) fact_ext) = Im(s11 +- s22) fext */
              double ip = (s[1] + s[7]) * fext, imn = (s[1] - s[7]) * fext;
              EX(0, 0, jj) = EX(0, 0, jj) + phi0_w * ip;
              EX(0, 1, jj) = EX(0, 1, jj) + phi0_w * imn;
              EX(1, 0, jj) = EX(1, 0, jj) + phi0_w * imn;
              EX(1, 1, jj) = EX(1, 1, jj) + phi0_w * ip;
            }
          }
          SM(0, ll, 0, jj, kk) = SM(0, ll, 0, jj, kk) + t11 * phi0_w;
          SM(0, ll, 1, jj, kk) = SM(0, ll, 1, jj, kk) + t12 * phi0_w;
          SM(1, ll, 0, jj, kk) = SM(1, ll, 0, jj, kk) + t21 * phi0_w;
          SM(1, ll, 1, jj, kk) = SM(1, ll, 1, jj, kk) + t22 * phi0_w;
        }
        emis11[ll + kk * NMU] = SM(0, ll, 0, jj, kk) * thet_w * 2.0 * PI_D;
        emis12[ll + kk * NMU] = SM(0, ll, 1, jj, kk) * thet_w * 2.0 * PI_D;
      }
    }
    double s11 = 0.0, s12 = 0.0;
    for (int i = 0; i < 2 * NMU; i++) {
      s11 += emis11[i];
      s12 += emis12[i];
    }
    emis[0 + 2 * jj] = EX(0, 0, jj) - s11;
    emis[1 + 2 * jj] = EX(0, 1, jj) - s12;
  }
#undef SM
#undef EX
  free(T.t);
  return 0;
}

/* ---- exported entry points (tests and the layer optics of pam_oracle.c) ---- */

/* The stored T matrix of one particle: returns nmax (0 on failure) and, when
 * t is not NULL and cap >= nmax, the 8 REAL*4 components as
 * t[c][m][n-1][nn-1] with strides (cap+1, cap, cap). */
int pam_oracle_tm_tmatrix(double lam, double mrr, double mri, double axi, double as_ratio,
                          float *tout, int cap) {
  tm_work *t = (tm_work *)calloc(1, sizeof(tm_work));
  tm_tmat T = {0, NULL};
  int st = tm_solve(t, lam, mrr, fabs(mri), axi, as_ratio, &T);
  free(t);
  if (st) return 0;
  int N = T.nmax;
  if (tout && cap >= N)
    for (int c = 0; c < 8; c++)
      for (int m = 0; m <= N; m++)
        for (int n = 1; n <= N; n++)
          for (int nn = 1; nn <= N; nn++)
            tout[(((size_t)c * (cap + 1) + m) * cap + (n - 1)) * cap + (nn - 1)] =
                *tm_at(&T, c, m, n, nn);
  free(T.t);
  return N;
}

/* calc_single_tmatrix for one particle (Fortran-order outputs as above) */
int pam_oracle_tm_single(double freq_hz, double mrr, double mri, double axi, double as_ratio,
                         double canting, double *smat, double *ext, double *emis, double *sback) {
  tm_work *t = (tm_work *)calloc(1, sizeof(tm_work));
  int st = tm_single(t, freq_hz, mrr, mri, axi, as_ratio, canting, smat, ext, emis, sback);
  free(t);
  return st;
}

/* calc_tmatrix (tmatrix.f90:14-444) over the size bins of one hydrometeor:
 * (mre, mim) the refractive index m = mre - i mim of scatProperties.f90:222,
 * phase = liq_ice (ice below 917 kg/m3 is mixed with air, eps_mix).
 * Outputs as ref_calc_tmatrix: smat(2,NMU,2,NMU,2), ext(2,2,NMU),
 * emis(2,NMU), sback(2,2), all sums over the bins of part * ndens * del_d. */
int pam_oracle_calc_tmatrix(double freq_hz, double mre, double mim, int phase, int nbins,
                            const double *dmax, const double *del_d, const double *ndens,
                            const double *density, const double *as_ratio,
                            const double *canting, double *smat, double *ext, double *emis,
                            double *sback) {
  const int ns = 2 * NMU * 2 * NMU * 2;
  memset(smat, 0, sizeof(double) * ns);
  memset(ext, 0, sizeof(double) * 4 * NMU);
  memset(emis, 0, sizeof(double) * 2 * NMU);
  if (sback) memset(sback, 0, sizeof(double) * 4);
  if (isnan(freq_hz) || freq_hz < 0.0 || isnan(mre) || isnan(mim) || nbins <= 0)
    return PAM_FATAL;
  for (int i = 0; i < nbins; i++)
    if (isnan(dmax[i]) || dmax[i] <= 0.0 || isnan(del_d[i]) || del_d[i] <= 0.0 ||
        isnan(ndens[i]) || ndens[i] < 0.0 || isnan(density[i]) || density[i] < 0.0 ||
        isnan(as_ratio[i]) || as_ratio[i] <= 0.0 || isnan(canting[i]) || canting[i] < 0.0)
      return PAM_FATAL;
  tm_work *t = (tm_work *)calloc(1, sizeof(tm_work));
  double *sp = (double *)malloc(sizeof(double) * (ns + 6 * NMU + 4));
  double *ep = sp + ns, *mp = ep + 4 * NMU, *bp = mp + 2 * NMU;
  int st = 0;
  for (int ir = 0; ir < nbins && st == 0; ir++) {
    double nd = ndens[ir], dd = del_d[ir];
    if (nd == 0.0) continue;
    if (!(density[ir] > 0.0)) {
      st = PAM_FATAL;
      break;
    }
    double re = mre, im = -mim; /* mMix = ref_index */
    if (phase == -1 && density[ir] != 917.0) {
      double r2, i2;
      pam_oracle_eps_mix(mre, -mim, density[ir], &r2, &i2);
      re = r2;
      im = i2;
    }
    /* mindex = conjg(mMix) */
    double mr = re, mi = -im;
    double axi = as_ratio[ir] <= 1 ? 0.5 * dmax[ir] * pow(as_ratio[ir], 1.0 / 3.0)
                                   : 0.5 * dmax[ir] / pow(as_ratio[ir], 2.0 / 3.0);
    memset(bp, 0, sizeof(double) * 4);
    st = tm_single(t, freq_hz, mr, mi, axi, as_ratio[ir], canting[ir], sp, ep, mp, bp);
    if (st) break;
    for (int i = 0; i < ns; i++) smat[i] = smat[i] + sp[i] * nd * dd;
    for (int i = 0; i < 4 * NMU; i++) ext[i] = ext[i] + ep[i] * nd * dd;
    for (int i = 0; i < 2 * NMU; i++) emis[i] = emis[i] + mp[i] * nd * dd;
    if (sback)
      for (int i = 0; i < 4; i++) sback[i] = sback[i] + bp[i] * nd * dd;
  }
  free(sp);
  free(t);
  return st ? PAM_FATAL : 0;
}
