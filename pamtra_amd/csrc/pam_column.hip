// pam_column.hip -- per-column preparation on the device:
//  * profile gap filling (vars_atmosphere.f90:671-901), output level search
//    (run_rt.f90:166-184)
//  * hydrometeor presence per layer (hydrometeor_extinction.f90:254-283)
//  * gas absorption, Rosenkranz 1998 (gasabs_module.f90:83-333)
//  * surface emissivity: TESSEM2 ocean (tessem2.f90) or a constant
//    (sfc_optics.f90:54-82)
// All are bandwidth-trivial elementwise kernels: one thread per column,
// per (column, frequency) or per (column, frequency, layer).
#include "pam_internal.h"

namespace pam {

// ---------------- Rosenkranz 1998 line tables (single-precision literals
// in the reference, gasabs_module.f90:141-189, 269-297) ----------------
__constant__ float c_o2_w300[40] = {1.63f, 1.646f, 1.468f, 1.449f, 1.382f, 1.360f, 1.319f,
  1.297f, 1.266f, 1.248f, 1.221f, 1.207f, 1.181f, 1.171f, 1.144f, 1.139f, 1.110f, 1.108f,
  1.079f, 1.078f, 1.05f, 1.05f, 1.02f, 1.02f, 1.0f, 1.0f, 0.97f, 0.97f, 0.94f, 0.94f, 0.92f,
  0.92f, 0.89f, 0.89f, 1.92f, 1.92f, 1.92f, 1.81f, 1.81f, 1.81f};
__constant__ float c_o2_y300[40] = {-0.0233f, 0.2408f, -0.3486f, 0.5227f, -0.5430f, 0.5877f,
  -0.3970f, 0.3237f, -0.1348f, 0.0311f, 0.0725f, -0.1663f, 0.2832f, -0.3629f, 0.3970f,
  -0.4599f, 0.4695f, -0.5199f, 0.5187f, -0.5597f, 0.5903f, -0.6246f, 0.6656f, -0.6942f,
  0.7086f, -0.7325f, 0.7348f, -0.7546f, 0.7702f, -0.7864f, 0.8083f, -0.8210f, 0.8439f,
  -0.8529f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
__constant__ float c_o2_v[40] = {0.0079f, -0.0978f, 0.0844f, -0.1273f, 0.0699f, -0.0776f,
  0.2309f, -0.2825f, 0.0436f, -0.0584f, 0.6056f, -0.6619f, 0.6451f, -0.6759f, 0.6547f,
  -0.6675f, 0.6135f, -0.6139f, 0.2952f, -0.2895f, 0.2654f, -0.2590f, 0.3750f, -0.3680f,
  0.5085f, -0.5002f, 0.6206f, -0.6091f, 0.6526f, -0.6393f, 0.6640f, -0.6475f, 0.6729f,
  -0.6545f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
__constant__ float c_o2_f[40] = {118.7503f, 56.2648f, 62.4863f, 58.4466f, 60.3061f, 59.5910f,
  59.1642f, 60.4348f, 58.3239f, 61.1506f, 57.6125f, 61.8002f, 56.9682f, 62.4112f, 56.3634f,
  62.9980f, 55.7838f, 63.5685f, 55.2214f, 64.1278f, 54.6712f, 64.6789f, 54.1300f, 65.2241f,
  53.5957f, 65.7648f, 53.0669f, 66.3021f, 52.5424f, 66.8368f, 52.0214f, 67.3696f, 51.5034f,
  67.9009f, 368.4984f, 424.7631f, 487.2494f, 715.3932f, 773.8397f, 834.1453f};
__constant__ float c_o2_s300[40] = {0.2936E-14f, 0.8079E-15f, 0.2480E-14f, 0.2228E-14f,
  0.3351E-14f, 0.3292E-14f, 0.3721E-14f, 0.3891E-14f, 0.3640E-14f, 0.4005E-14f, 0.3227E-14f,
  0.3715E-14f, 0.2627E-14f, 0.3156E-14f, 0.1982E-14f, 0.2477E-14f, 0.1391E-14f, 0.1808E-14f,
  0.9124E-15f, 0.1230E-14f, 0.5603E-15f, 0.7842E-15f, 0.3228E-15f, 0.4689E-15f, 0.1748E-15f,
  0.2632E-15f, 0.8898E-16f, 0.1389E-15f, 0.4264E-16f, 0.6899E-16f, 0.1924E-16f, 0.3229E-16f,
  0.8191E-17f, 0.1423E-16f, 0.6460E-15f, 0.7047E-14f, 0.3011E-14f, 0.1826E-14f, 0.1152E-13f,
  0.3971E-14f};
__constant__ float c_o2_be[40] = {0.009f, 0.015f, 0.083f, 0.084f, 0.212f, 0.212f, 0.391f,
  0.391f, 0.626f, 0.626f, 0.915f, 0.915f, 1.26f, 1.26f, 1.66f, 1.665f, 2.119f, 2.115f, 2.624f,
  2.625f, 3.194f, 3.194f, 3.814f, 3.814f, 4.484f, 4.484f, 5.224f, 5.224f, 6.004f, 6.004f,
  6.844f, 6.844f, 7.744f, 7.744f, 0.048f, 0.044f, 0.049f, 0.145f, 0.141f, 0.145f};
// water vapour lines: fl, s1, b2, w3, x, ws, xs
__constant__ float c_h2o[7][15] = {
    {22.2351f, 183.3101f, 321.2256f, 325.1529f, 380.1974f, 439.1508f, 443.0183f, 448.0011f,
     470.8890f, 474.6891f, 488.4911f, 556.9360f, 620.7008f, 752.0332f, 916.1712f},
    {0.1310E-13f, 0.2273E-11f, 0.8036E-13f, 0.2694E-11f, 0.2438E-10f, 0.2179E-11f, 0.4624E-12f,
     0.2562E-10f, 0.8369E-12f, 0.3263E-11f, 0.6659E-12f, 0.1531E-08f, 0.1707E-10f, 0.1011E-08f,
     0.4227E-10f},
    {2.144f, .668f, 6.179f, 1.541f, 1.048f, 3.595f, 5.048f, 1.405f, 3.597f, 2.379f, 2.852f,
     .159f, 2.391f, .396f, 1.441f},
    {0.002656f, 0.00281f, 0.0023f, 0.00278f, 0.00287f, 0.0021f, 0.00186f, 0.00263f, 0.00215f,
     0.00236f, 0.0026f, 0.00321f, 0.00244f, 0.00306f, 0.00267f},
    {0.69f, 0.64f, 0.67f, 0.68f, 0.54f, 0.63f, 0.60f, 0.66f, 0.66f, 0.65f, 0.69f, 0.69f, 0.71f,
     0.68f, 0.70f},
    {0.0127488f, 0.01491f, 0.0108f, 0.0135f, 0.01541f, 0.0090f, 0.00788f, 0.01275f, 0.00983f,
     0.01095f, 0.01313f, 0.01320f, 0.01140f, 0.01253f, 0.01275f},
    {0.61f, 0.85f, 0.54f, 0.74f, 0.89f, 0.52f, 0.50f, 0.67f, 0.65f, 0.64f, 0.72f, 1.0f, 0.68f,
     0.84f, 0.78f}};

// TESSEM2 network weights (tessem2.f90:40-94), single-precision literals
__constant__ float c_ts_b1[2][15] = {
    {-4.425637f, -9.833792f, 1.338131f, -1.688713f, 0.121844f, -0.020075f, 0.983048f,
     -5.365536f, 0.759206f, 1.705090f, -0.861439f, 2.639012f, 4.438483f, -1.376479f, 0.920075f},
    {-0.492534f, 3.549479f, 1.209267f, -1.201308f, 0.657236f, 0.522424f, 0.439041f, 1.070375f,
     0.917859f, 2.999629f, 0.150532f, -0.225454f, -1.418973f, -5.616032f, 3.077098f}};
__constant__ float c_ts_b2[2] = {-3.126669f, -3.191476f};
// first-layer weights, [pol][node + 15 * input]
__constant__ float c_ts_w1[2][75] = {
    {0.988378f, -0.053936f, -1.166637f, 0.213664f, -0.264162f, -0.946456f, -0.084722f,
     -2.169090f, 1.038926f, 1.814769f, -0.588050f, 1.553885f, 3.672554f, 0.198168f, 0.520655f,
     -0.865033f, 6.091850f, 0.514930f, 1.203841f, 1.349278f, -0.875449f, -1.582926f, 1.683341f,
     1.679280f, 0.533539f, 1.047984f, 0.534832f, 0.429764f, 1.978477f, -0.982343f, 0.024583f,
     -2.832310f, -0.249698f, -0.264623f, -0.385522f, 0.197084f, 0.071463f, -1.025112f,
     -0.365062f, 0.099878f, 0.308190f, -0.034068f, 0.011647f, 0.158408f, -0.512171f, -1.243063f,
     -0.008161f, 0.323026f, -0.040724f, 0.526443f, -1.044800f, 0.035568f, 0.057517f, -0.559605f,
     -0.086835f, 0.058220f, 0.284328f, 0.235683f, -0.050183f, -0.074034f, 2.514267f, 0.002472f,
     -0.249759f, -0.017623f, 2.265044f, 1.006211f, -0.000011f, -0.010390f, 0.285759f,
     -0.024043f, -0.010873f, -0.063544f, -0.043210f, -0.000275f, 0.017342f},
    {0.914705f, -0.244028f, 0.217031f, 0.022331f, -0.757337f, -0.514908f, -0.682990f, 1.319635f,
     1.052532f, 0.327665f, -0.197895f, 0.529253f, -1.846079f, -5.075406f, 3.362876f, -0.907370f,
     -1.694778f, -0.628096f, 2.230164f, 0.003549f, -0.290958f, -0.013354f, 0.122967f, 0.001593f,
     -0.061660f, 1.609950f, -0.056635f, 1.059906f, -0.224954f, 0.942301f, 1.806560f, 0.821685f,
     -1.057938f, 0.082932f, 0.045666f, 0.119414f, 0.032902f, -0.178454f, -0.058426f, -0.098223f,
     0.291156f, -0.036004f, -0.524861f, 0.089880f, 0.268074f, -0.237687f, 0.040081f, -0.019350f,
     0.028382f, 0.831685f, -1.168739f, 0.715660f, 0.498660f, -0.716106f, 0.916730f, 0.060666f,
     -0.605959f, 0.214584f, -0.233417f, -0.207224f, 0.013675f, -0.000319f, 0.008815f, -0.004765f,
     0.823811f, -0.244888f, 0.265740f, 1.596743f, -2.614260f, -2.705295f, -0.003804f, 0.105391f,
     -0.008504f, 0.003713f, -0.018812f}};
__constant__ float c_ts_w2[2][15] = {
    {-0.105095f, -5.132794f, -0.068984f, -1.133355f, -0.016834f, 0.021879f, -1.791876f,
     2.727342f, 0.070811f, 0.556869f, -1.172738f, -1.725109f, 1.650515f, -1.123763f, -0.929365f},
    {0.019943f, 2.713515f, -0.147582f, -0.305444f, -0.275328f, -0.054163f, 0.628523f, 0.028561f,
     0.044241f, 0.034885f, -0.168618f, 0.580301f, -0.157557f, -0.586725f, 0.179877f}};
__constant__ float c_ts_xmin[5] = {10.0f, 0.0f, 0.0f, 273.149990f, 0.0f};
__constant__ float c_ts_xmax[5] = {700.0f, 88.0f, 24.0f, 309.149994f, 40.0f};
__constant__ float c_ts_ymin[2] = {0.272829f, 0.018036f};
__constant__ float c_ts_ymax[2] = {0.997013f, 0.884545f};

// Goff-Gratch saturation pressure over water, hPa (e_sat_gg_water.f90:1-21)
__device__ double esat_gg(double T) {
  double a = sp(373.16) / T;
  double s = -sp(7.90298) * (a - 1.0) + sp(5.02808) * log10(a) -
             sp(1.3816e-7) * (pow(10.0, sp(11.344) * (1.0 - T / sp(373.16))) - 1.0) +
             sp(8.1328e-3) * (pow(10.0, -sp(3.49149) * (a - 1.0)) - 1.0);
  return sp(1013.246) * pow(10.0, s);
}

__device__ __forceinline__ double ld_or_nan(const double *a, size_t i) {
  return a ? a[i] : __builtin_nan("");
}

// One thread per column: level/layer gap filling and output levels.
__global__ void k_fill_columns(Dims d, Profile p, Work w) {
  int cl = blockIdx.x * blockDim.x + threadIdx.x;
  if (cl >= d.ncol_chunk) return;
  int col = d.col0 + cl;
  const int ML = d.ML;
  int L = p.nlyrs[col];
  w.nl[col] = L;
  int32_t st = 0;
  if (L <= 0 || L > ML) {
    w.col_status[col] = E_NLYRS | PAM_FATAL;
    return;
  }
  size_t ol = (size_t)col * (ML + 1), oy = (size_t)col * ML;
  double *hl = w.hgt_lev + ol, *tl = w.temp_lev + ol;
  double *t = w.temp + oy, *pr = w.press + oy;
  // level and layer inputs, NaN = missing; the filled copies live in the
  // workspace, relhum and layer heights only in registers-backed scratch
  double *rl = w.rho_v + oy;  // scratch for relhum layers (overwritten below)
  double *h = w.q_hum + oy;   // scratch for layer heights (overwritten below)
  double pl0 = 0.0, plL = 0.0;
  (void)pl0;
  (void)plL;
  for (int z = 0; z <= L; z++) {
    hl[z] = ld_or_nan(p.hgt_lev, ol + z);
    tl[z] = ld_or_nan(p.temp_lev, ol + z);
  }
  for (int z = 0; z < L; z++) {
    h[z] = ld_or_nan(p.hgt, oy + z);
    t[z] = ld_or_nan(p.temp, oy + z);
    pr[z] = ld_or_nan(p.press, oy + z);
    rl[z] = ld_or_nan(p.relhum, oy + z);
  }
  for (int z = 0; z < L; z++)
    if (isnan(h[z])) h[z] = 0.5 * (hl[z] + hl[z + 1]);
  // press_lev / relhum_lev are only needed where a layer value is missing
  auto plev = [&](int z) -> double {
    double v = ld_or_nan(p.press_lev, ol + z);
    if (!isnan(v)) return v;
    if (L < 2) return v;
    if (z == 0)
      return exp(log(pr[0]) - (log(pr[1]) - log(pr[0])) / (h[1] - h[0]) * (h[0] - hl[0]));
    if (z == L)
      return exp(log(pr[z - 1]) + (log(pr[z - 1]) - log(pr[z - 2])) / (h[z - 1] - h[z - 2]) *
                                      (hl[z] - h[z - 1]));
    return exp(0.5 * (log(pr[z]) + log(pr[z - 1])));
  };
  auto rlev = [&](int z) -> double {
    double v = ld_or_nan(p.relhum_lev, ol + z);
    if (!isnan(v)) return v;
    if (L < 2) return v;
    if (z == 0) return rl[0] - (rl[1] - rl[0]) / (h[1] - h[0]) * (h[0] - hl[0]);
    if (z == L)
      return rl[z - 1] + (rl[z - 1] - rl[z - 2]) / (h[z - 1] - h[z - 2]) * (hl[z] - h[z - 1]);
    return 0.5 * (rl[z] + rl[z - 1]);
  };
  if (L >= 2) {
    if (isnan(hl[0])) hl[0] = h[0] - (h[1] - h[0]) * 0.5;
    if (isnan(tl[0])) tl[0] = t[0] - (t[1] - t[0]) / (h[1] - h[0]) * (h[0] - hl[0]);
  }
  for (int z = 1; z < L; z++) {
    if (isnan(tl[z])) tl[z] = 0.5 * (t[z] + t[z - 1]);
    if (isnan(hl[z])) hl[z] = 0.5 * (h[z] + h[z - 1]);
  }
  if (L >= 2) {
    int z = L;
    if (isnan(hl[z])) hl[z] = h[z - 1] + (h[z - 1] - h[z - 2]) * 0.5;
    if (isnan(tl[z]))
      tl[z] = t[z - 1] + (t[z - 1] - t[z - 2]) / (h[z - 1] - h[z - 2]) * (hl[z] - h[z - 1]);
  }
  // layer values from levels where missing
  for (int z = 0; z < L; z++) {
    if (isnan(t[z])) t[z] = 0.5 * (tl[z] + tl[z + 1]);
    if (isnan(rl[z])) rl[z] = 0.5 * (rlev(z) + rlev(z + 1));
    if (isnan(pr[z])) pr[z] = exp(0.5 * (log(plev(z)) + log(plev(z + 1))));
  }
  for (int z = 0; z <= L; z++)
    if (isnan(tl[z]) || isnan(hl[z])) st = E_NAN_LEVEL | PAM_FATAL;
  for (int z = 0; z < L; z++) {
    double vp = rl[z] * esat_gg(t[z]);
    double rho_v = vp / (t[z] * R_V);
    double q = R_V / R_D * vp / (pr[z] - (1.0 - R_V / R_D) * vp);
    w.dz[oy + z] = hl[z + 1] - hl[z];
    w.hgt[oy + z] = h[z];
    rl[z] = rho_v;  // w.rho_v
    h[z] = q;       // w.q_hum
    if (isnan(t[z]) || isnan(pr[z]) || isnan(rho_v) || isnan(q)) st = E_NAN_LAYER | PAM_FATAL;
  }
  if (!(hl[L] > hl[0])) st = E_HGT_ORDER | PAM_FATAL;
  double gt = p.groundtemp ? p.groundtemp[col] : __builtin_nan("");
  w.groundtemp[col] = isnan(gt) ? tl[0] : gt;
  // output levels
  for (int io = 0; io < d.nout; io++) {
    double oh = p.obs_height ? p.obs_height[(size_t)col * d.nout + io]
                             : (io == 0 ? 833000. : 0.);
    if (isnan(oh)) oh = (io == 0) ? 833000. : 0.;
    int o;
    if (oh >= hl[L])
      o = 1;
    else if (oh <= hl[0])
      o = L + 1;
    else {
      o = L + 1;
      for (int nz = 1; nz <= L; nz++)
        if (hl[nz] >= oh) {
          o = (fabs(hl[nz] - oh) <= fabs(hl[nz - 1] - oh)) ? L - nz + 1 : L - nz + 2;
          break;
        }
    }
    w.outlev[(size_t)col * d.nout + io] = o;
  }
  w.col_status[col] = st;
}

__device__ __forceinline__ double q2abs(double spec, double t, double p, double qv,
                                        double qall) {
  return spec * p / (R_D * (1.0 + (R_V / R_D - 1.0) * qv - qall) * t);
}

// One thread per column: which hydrometeors are present in each layer
// (hydrometeor_extinction.f90:236-285) and the number of scattering layers.
__global__ void k_layer_flags(Dims d, Settings s, Profile p, const pam_hydro_desc *hyd,
                              Hydro4D h4, Work w) {
  int cl = blockIdx.x * blockDim.x + threadIdx.x;
  if (cl >= d.ncol_chunk) return;
  int col = d.col0 + cl;
  int L = w.nl[col];
  int n = 0;
  if (w.col_status[col] & 0xff) {
    for (int z = 0; z < d.ML; z++) w.hmask[(size_t)col * d.ML + z] = 0;
    w.col_nitems[col] = 0;
    return;
  }
  for (int z = 0; z < d.ML; z++) {
    uint32_t mask = 0;
    if (z < L) {
      size_t b4 = ((size_t)col * d.ML + z) * d.nh;
      double qsum = 0.0;
      bool qnan = false;
      for (int hh = 0; hh < d.nh; hh++) {
        double q = p.hydro_q ? p.hydro_q[b4 + hh] : __builtin_nan("");
        if (isnan(q)) qnan = true;
        qsum += q;
      }
      size_t oy = (size_t)col * d.ML + z;
      double T = w.temp[oy], P = w.press[oy], qv = w.q_hum[oy];
      double qall = (s.incl_hydro_rho && !qnan) ? qsum : 0.0;
      for (int hh = 0; hh < d.nh && h4.fs_nbin > 0; hh++) {
        // full spectrum: present when its bins hold particles or the
        // threshold is not positive (hydrometeor_extinction.f90:141-148)
        const double *n = h4.fs_n + (b4 + hh) * (size_t)h4.fs_nbin;
        double nsum = 0.0;
        for (int i = 0; i < h4.fs_nbin; i++) nsum += n[i];
        if (nsum > 0.0 || s.threshold <= 0.0) mask |= (1u << hh);
      }
      for (int hh = 0; hh < d.nh && h4.fs_nbin == 0; hh++) {
        const pam_hydro_desc &hd = hyd[hh];
        if (hd.dist_name == PAM_DIST_DISABLED) continue;
        double qin = p.hydro_q ? p.hydro_q[b4 + hh] : __builtin_nan("");
        double nin = p.hydro_n ? p.hydro_n[b4 + hh] : __builtin_nan("");
        double q_h = q2abs(qin, T, P, qv, qall);
        double n_tot = q2abs(nin, T, P, qv, qall);
        double r_eff = p.hydro_reff ? p.hydro_reff[b4 + hh] : __builtin_nan("");
        if (hd.moment_in == 13 && (q_h < s.threshold || n_tot <= 1.)) continue;
        if ((q_h < s.threshold || isnan(q_h)) && (n_tot <= 0 || isnan(n_tot)) &&
            (r_eff <= 0 || isnan(r_eff)) && hd.moment_in != 0)
          continue;
        mask |= (1u << hh);
      }
    }
    w.hmask[(size_t)col * d.ML + z] = mask;
    if (mask) n++;
  }
  w.col_nitems[col] = n;
}

// Rosenkranz 1998 absorption, Np/m (rosen98_gasabs.f90 + get_gasabs.f90)
__device__ double gasabs_r98(double f, double tk, double rhowv, double pres, bool &ok) {
  ok = !(f <= 0.0 || f > 800.0 || tk <= 100.0 || pres <= 0.0 || pres > 1.2e5);
  double pmb = pres / 100.0;
  double vapden = rhowv * 1000.0;
  double th = 300.0 / tk;
  double absn2 = 6.4e-14 * pmb * pmb * f * f * pow(th, 3.55);
  // o2abs
  double th1 = th - 1.0;
  double b = pow(th, sp(0.8));
  double preswv = vapden * tk / 217.0;
  double presda = pmb - preswv;
  double den = 0.001 * (presda * b + 1.1 * preswv * th);
  double dfnr = sp(0.56) * den;
  double sum = 1.6e-17 * f * f * dfnr / (th * (f * f + dfnr * dfnr));
  for (int k = 0; k < 40; k++) {
    double fk = (double)c_o2_f[k];
    double df = (double)c_o2_w300[k] * den;
    double y = 0.001 * pmb * b * ((double)c_o2_y300[k] + (double)c_o2_v[k] * th1);
    double str = (double)c_o2_s300[k] * exp(-(double)c_o2_be[k] * th1);
    double sf1 = (df + (f - fk) * y) / ((f - fk) * (f - fk) + df * df);
    double sf2 = (df - (f + fk) * y) / ((f + fk) * (f + fk) + df * df);
    double r = f / fk;
    sum = sum + str * (sf1 + sf2) * (r * r);
  }
  double o2 = 0.5034e12 * sum * presda * (th * th * th) / PI;
  // abh2o
  double h2o = 0.0;
  if (vapden > 0.0) {
    double pvap = vapden * tk / 217.0;
    double pda = pmb - pvap;
    double dens = 3.335e16 * vapden;
    double ti = 300.0 / tk;
    double ti2 = pow(ti, 2.5);
    double con = (5.43e-10 * 1.105 * pda * (ti * ti * ti) + 1.8e-8 * 0.79 * pvap * pow(ti, 7.5)) *
                 pvap * f * f;
    double s2 = 0.0;
    for (int i = 0; i < 15; i++) {
      double fl = (double)c_h2o[0][i];
      double width = (double)c_h2o[3][i] * pda * pow(ti, (double)c_h2o[4][i]) +
                     (double)c_h2o[5][i] * pvap * pow(ti, (double)c_h2o[6][i]);
      double wsq = width * width;
      double ss = (double)c_h2o[1][i] * ti2 * exp((double)c_h2o[2][i] * (1.0 - ti));
      double df0 = f - fl, df1 = f + fl;
      double base = width / (562500.0 + wsq);
      double res = 0.0;
      if (fabs(df0) < 750.0) res = res + width / (df0 * df0 + wsq) - base;
      if (fabs(df1) < 750.0) res = res + width / (df1 * df1 + wsq) - base;
      double r = f / fl;
      s2 = s2 + ss * res * (r * r);
    }
    h2o = 0.3183e-4 * dens * s2 + con;
  }
  return (absn2 + o2 + h2o) / 1.0e3;
}

// One thread per (column, frequency, layer).
__global__ void k_gas(Dims d, Settings s, const double *freqs, Work w) {
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t n = (size_t)d.ncol_chunk * d.nf * d.ML;
  if (idx >= n) return;
  int z = idx % d.ML;
  int f = (idx / d.ML) % d.nf;
  int col = d.col0 + (int)(idx / ((size_t)d.ML * d.nf));
  size_t out = ((size_t)col * d.nf + f) * d.ML + z;
  if (z >= w.nl[col] || (w.col_status[col] & 0xff)) {
    w.kgas[out] = 0.0;
    return;
  }
  size_t oy = (size_t)col * d.ML + z;
  double k = 0.0;
  if (s.lgas) {
    bool ok;
    k = gasabs_r98(freqs[f], w.temp[oy], w.rho_v[oy], w.press[oy], ok);
    if (!ok) set_err(&w.cf_status[(size_t)col * d.nf + f], E_GAS);
  }
  w.kgas[out] = k;
}

// TESSEM2 ocean emissivity (tessem2.f90:1-137)
__device__ void tessem2(double freq, double zen, double wind, double temp, double sal,
                        double *emis) {
  double x[5] = {freq, zen, wind, temp, sal};
  for (int pol = 0; pol < 2; pol++) {
    double nx[5];
    for (int i = 0; i < 5; i++)
      nx[i] = -1.0 + (x[i] - (double)c_ts_xmin[i]) /
                         ((double)c_ts_xmax[i] - (double)c_ts_xmin[i]) * 2.0;
    double ny = (double)c_ts_b2[pol];
    for (int i = 0; i < 15; i++) {
      double t = (double)c_ts_b1[pol][i];
      for (int j = 0; j < 5; j++) t = t + (double)c_ts_w1[pol][i + 15 * j] * nx[j];
      double tr = 2.0 / (1.0 + exp(-2.0 * t)) - 1.0;
      ny = ny + (double)c_ts_w2[pol][i] * tr;
    }
    emis[pol] = (double)c_ts_ymin[pol] +
                (ny + 1.0) / 2.0 * ((double)c_ts_ymax[pol] - (double)c_ts_ymin[pol]);
  }
}

// FASTEM5 ocean emissivity as PAMTRA calls it (ocean_sfc_optics.f90:68-107,
// fastemx.f90:120-301): Liu double-Debye sea water (liu.f90), Fresnel
// (fresnel.f90), foam (foam_utility_module.f90), large- and small-scale
// roughness (large_/small_scale_correction_module.f90); no azimuth term and
// no reflection correction.  zen in degrees, temp in K, sal in psu, f in GHz.
__constant__ double c_fe_lsc[2][6][3] = {
    {{-5.994667E-002, 9.341346E-004, -9.56611E-007},
     {8.360313E-002, -1.085991E-003, 6.735338E-007},
     {-2.617296E-002, 2.864495E-004, -1.429979E-007},
     {-5.265879E-004, 6.880275E-005, -2.916657E-007},
     {-1.671574E-005, 1.086405E-006, -3.632227E-009},
     {1.16194E-004, -6.349418E-005, 2.466556E-007}},
    {{-2.431811E-002, -1.03181E-003, 4.519513E-006},
     {2.868236E-002, 1.186478E-003, -5.257096E-006},
     {-7.93339E-003, -2.422303E-004, 1.089605E-006},
     {-1.083452E-003, -1.788509E-005, 5.464239E-009},
     {-3.855673E-005, 9.360072E-007, -2.639362E-009},
     {1.101309E-003, 3.599147E-005, -1.043146E-007}}};
__constant__ double c_fe_ssc[8] = {-5.020848e-006,  2.3297951e-008, 4.6625726e-008,
                                   -1.9765665e-009, -7.0469823e-004, 7.5061193e-004,
                                   9.8103876e-004,  1.54895e-004};

__device__ cd liu_permittivity(double temp, double sal, double f) {
  const double mu0 = PI * 4.0e-07;                 // constants.f90:34
  const double eps0 = 1.0 / (mu0 * C_LIGHT * C_LIGHT);
  const double tau0 = 2.0 * PI * eps0 * 1.0e9;
  const double t = temp - 273.15, s = sal;
  const double einf = 3.8 + t * 2.48033e-02;
  double es = 87.9181727 + t * (-4.031592248e-01 + t * (9.493088010e-04 + t * -1.930858348e-06));
  double e1 = 5.723 + t * (2.2379e-02 + t * -7.1237e-04);
  double tau1 = 1.124465e-01 + t * (-3.9815727e-03 + t * (8.113381e-05 + t * -7.1824242e-07));
  double tau2 =
      3.049979018e-03 + t * (-3.010041629e-05 + t * (4.811910733e-06 + t * -4.259775841e-08));
  double ie = 0.0;
  if (s > 0.0) {
    const double delta = 25.0 - t;
    const double beta = 2.033e-02 + delta * (1.266e-04 + delta * 2.464e-06) +
                        (-1.849e-05 + delta * (2.551e-07 + delta * -2.551e-08)) * s;
    const double alpha25 =
        s * (1.82521e-01 + s * (-1.46192e-03 + s * (2.09324e-05 + s * -1.28205e-07)));
    ie = -(alpha25 * exp(-delta * beta)) / (f * tau0);
    es = es * (1.0 + s * (-2.697e-03 + s * -7.3e-06 + t * -8.9e-06));
    e1 = e1 * (1.0 + s * (-6.28908e-03 + s * 1.76032e-04 + t * -9.22144e-05));
    tau1 = tau1 * (1.0 + s * (-2.39357e-03 + t * (3.1353e-05 + t * -2.52477e-07)));
    tau2 = tau2 * (1.0 + s * (1.49e-01 + t * -8.8e-04 + (s * s) * -1.05e-04));
  }
  const double f1 = f * tau1, f2 = f * tau2;
  const double d1 = es - e1, d2 = e1 - einf;
  const double re = einf + d1 / (1.0 + f1 * f1) + d2 / (1.0 + f2 * f2);
  ie = -ie + d1 * f1 / (1.0 + f1 * f1) + d2 * f2 / (1.0 + f2 * f2);
  return cd{re, -ie};
}

__device__ void fastem(double f, double zen, double temp, double sal, double wind, double *emis) {
  const double cz = cos(zen * (PI / 180.0));
  const cd eps = liu_permittivity(temp, sal, f);
  const cd z1 = csqrt_(eps - 1.0 + cz * cz), z2 = eps * cz;
  const cd zh = cdiv(cd{cz, 0.0} - z1, cd{cz, 0.0} + z1), zv = cdiv(z2 - z1, z2 + z1);
  const double rvf = zv.re * zv.re + zv.im * zv.im, rhf = zh.re * zh.re + zh.im * zh.im;
  const double fh = 1.0 + zen * (-1.748e-3 + zen * (-7.336e-5 + zen * 1.044e-7));
  const double ff = 0.40 * exp(-5.0e-2 * f);
  const double rv_foam = (1.0 - 0.93) * ff, rh_foam = (1.0 - fh * 0.93) * ff;
  const double cover = wind < 0.0 ? 0.0 : 1.95e-5 * pow(wind, 2.55);
  const double sec = 1.0 / cz;
  double rl[2];
  for (int pol = 0; pol < 2; pol++) {
    double z[6];
    for (int i = 0; i < 6; i++)
      z[i] = c_fe_lsc[pol][i][0] + f * (c_fe_lsc[pol][i][1] + f * c_fe_lsc[pol][i][2]);
    rl[pol] = z[0] + z[1] * sec + z[2] * (sec * sec) + z[3] * wind + z[4] * (wind * wind) +
              z[5] * wind * sec;
  }
  const double fl = f < 1.4 ? 1.4 : (f > 200.0 ? 200.0 : f);
  const double wl = wind < 0.3 ? 0.3 : (wind > 35.0 ? 35.0 : wind);
  const double f2 = fl * fl, w2 = wl * wl;
  const double y = c_fe_ssc[0] * wl * fl + c_fe_ssc[1] * wl * f2 + c_fe_ssc[2] * w2 * fl +
                   c_fe_ssc[3] * w2 * f2 + c_fe_ssc[4] * w2 / fl + c_fe_ssc[5] * w2 / f2 +
                   c_fe_ssc[6] * wl + c_fe_ssc[7] * w2;
  const double fsmall = exp(-y * (cz * cz));
  double rv = rvf, rh = rhf;
  if (zen < 70.0) {
    rv = rvf * fsmall - rl[0];
    rh = rhf * fsmall - rl[1];
  }
  emis[0] = 1.0 - (1.0 - cover) * rv - cover * rv_foam;
  emis[1] = 1.0 - (1.0 - cover) * rh - cover * rh_foam;
}

// One thread per (column, frequency): surface emissivity/reflectivity.
__global__ void k_surface(Dims d, Settings s, Profile p, const double *freqs, Work w,
                          double *emis_out) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= d.ncol_chunk * d.nf) return;
  int f = idx % d.nf;
  int col = d.col0 + idx / d.nf;
  int stype = p.sfc_type ? p.sfc_type[col] : -9999;
  int smodel = p.sfc_model ? p.sfc_model[col] : -9999;
  double *E = w.sfc_emis + ((size_t)col * d.nf + f) * 2 * NMU;
  double *R = w.sfc_refl + ((size_t)col * d.nf + f) * 2 * NMU;
  if (stype == 0) {
    // sfc_model -9999/0: TESSEM2; 1: FASTEM (sfc_optics.f90:55-58)
    if (!(smodel == 0 || smodel == -9999 || smodel == 1)) {
      set_err(&w.cf_status[(size_t)col * d.nf + f], E_SURFACE);
      return;
    }
    double wu = p.wind10u ? p.wind10u[col] : 0.0, wv = p.wind10v ? p.wind10v[col] : 0.0;
    double ws = sqrt(wu * wu + wv * wv);
    double sal = p.sfc_salinity ? p.sfc_salinity[col] : 33.0;
    for (int i = 0; i < NMU; i++) {
      double zen = acos(w.mu[i]) * (180.0 / PI);
      double e[2];
      if (smodel == 1)
        fastem(freqs[f], zen, w.groundtemp[col], sal, ws, e);
      else
        tessem2(freqs[f], zen, ws, w.groundtemp[col], sal, e);
      double r0 = 1.0 - e[0], r1 = 1.0 - e[1];
      if (i == NMU - 1) {
        double em = (e[0] + e[1]) / 2.0, rm = (r0 + r1) / 2.0;
        e[0] = e[1] = em;
        r0 = r1 = rm;
      }
      E[i] = e[0];
      E[NMU + i] = e[1];
      R[i] = r0;
      R[NMU + i] = r1;
    }
  } else if (stype == 1) {
    set_err(&w.cf_status[(size_t)col * d.nf + f], E_SURFACE);
    return;
  } else {
    double em = p.sfc_emissivity ? p.sfc_emissivity[col] : s.emissivity;
    for (int i = 0; i < 2 * NMU; i++) {
      E[i] = em;
      R[i] = 1.0 - em;
    }
  }
  if (emis_out)
    for (int k = 0; k < 2; k++)
      for (int i = 0; i < NMU; i++)
        emis_out[(((size_t)col * 2 + k) * d.nf + f) * NMU + i] = E[k * NMU + i];
}

// One thread per column: item index of each scattering layer of the chunk.
__global__ void k_items(Dims d, Work w) {
  int cl = blockIdx.x * blockDim.x + threadIdx.x;
  if (cl >= d.ncol_chunk) return;
  int col = d.col0 + cl;
  int it = w.col_item0[col];
  for (int z = 0; z < d.ML; z++) {
    size_t o = (size_t)col * d.ML + z;
    if (w.hmask[o]) {
      w.layer_item[o] = it;
      int li = it - d.item0;
      w.item_col[li] = col;
      w.item_layer[li] = z;
      it++;
    } else {
      w.layer_item[o] = -1;
    }
  }
}

static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

void launch_columns(const Dims &d, const Settings &s, const Profile &p, const pam_hydro_desc *hyd,
                    const Hydro4D &h4, const double *freqs, Work &w, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_columns, dim3(nblk(d.ncol_chunk, 128)), dim3(128), 0, st, d, p, w);
  hipLaunchKernelGGL(k_layer_flags, dim3(nblk(d.ncol_chunk, 128)), dim3(128), 0, st, d, s, p, hyd,
                     h4, w);
  size_t ng = (size_t)d.ncol_chunk * d.nf * d.ML;
  hipLaunchKernelGGL(k_gas, dim3(nblk(ng, 256)), dim3(256), 0, st, d, s, freqs, w);
}

void launch_surface(const Dims &d, const Settings &s, const Profile &p, const double *freqs,
                    Work &w, double *emis_out, hipStream_t st) {
  hipLaunchKernelGGL(k_surface, dim3(nblk((size_t)d.ncol_chunk * d.nf, 128)), dim3(128), 0, st, d,
                     s, p, freqs, w, emis_out);
}

void launch_items(const Dims &d, Work &w, hipStream_t st) {
  hipLaunchKernelGGL(k_items, dim3(nblk(d.ncol_chunk, 128)), dim3(128), 0, st, d, w);
}

// The radar simulator in its "simple" mode (radar_simulator.f90:94-170) for
// one (column, channel) of the chunk, with the attenuation run_rt stores
// (run_rt.f90:118-123): one-way gas and hydrometeor attenuation [dB] per
// layer, and for each layer with hydrometeors (a scattering unit)
// Ze = 10 log10(1e18 back lambda^4 / (|K|^2 pi^5) / PIA) [dBz], the
// fill value -9999 when back = 0.  hydrometeor_extinction visits the layers
// bottom-up, so a top-down PIA sees the hydrometeor attenuation above the
// layer as not yet set (0); a bottom-up PIA sums the layers below.
__global__ void k_radar(Dims d, Settings s, const double *freqs, Work w, double *ze,
                        double *att_hydro, double *att_atmo, double *radar_hgt) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)d.ncol_chunk * d.nf) return;
  const int f = (int)(idx % d.nf), col = d.col0 + (int)(idx / d.nf);
  const int ML = d.ML, np = s.radar_npol;
  const int L = min(max(w.nl[col], 0), ML);
  const bool col_ok = (w.col_status[col] & 0xff) < PAM_FATAL;
  const bool ok = col_ok && (w.cf_status[(size_t)col * d.nf + f] & 0xff) < PAM_FATAL;
  const double wl = C_LIGHT / (freqs[f] * 1.e9);
  const double k2 = w.radar_k2 ? w.radar_k2[f] : 0.93;
  const double zconst = 1.e18 * (1.0 / (k2 * pow(PI, 5)));
  const double wl4 = pow(wl, 4);
  const size_t oy = (size_t)col * ML;
  const double *kg = w.kgas + ((size_t)col * d.nf + f) * ML;
  auto out = [&](int z) { return ((size_t)col * ML + z) * d.nf + f; };
  // fill values above the column, and everything of a failed unit
  for (int z = 0; z < ML; z++) {
    const bool in = z < L;
    att_atmo[out(z)] = !in ? 0.0 : (ok ? 10.0 * log10(exp(kg[z] * w.dz[oy + z])) : __builtin_nan(""));
    att_hydro[out(z)] = !in ? 0.0 : (ok ? 0.0 : __builtin_nan(""));
    for (int ip = 0; ip < np; ip++)
      ze[out(z) * np + ip] = (!in || ok) ? -9999.0 : __builtin_nan("");
    if (f == 0) radar_hgt[oy + z] = !in ? -9999.0 : (col_ok ? w.hgt[oy + z] : __builtin_nan(""));
  }
  if (!ok) return;
  // hydrometeor attenuation of every scattering layer
  for (int z = 0; z < L; z++) {
    const int item = w.layer_item[oy + z];
    if (item < 0) continue;
    const size_t itf = (size_t)(item - d.item0) * d.nf + f;
    att_hydro[out(z)] = 10.0 * log10(exp(w.kext[itf] * w.dz[oy + z]));
  }
  // reflectivities; the running sums follow the order of the reference's
  // sums over the layers above (top-down) or below (bottom-up)
  const bool td = s.radar_att == 1;
  double sh = 0.0, sa = 0.0;
  for (int i = 0; i < L; i++) {
    const int z = td ? L - 1 - i : i;
    const int item = w.layer_item[oy + z];
    if (item >= 0) {
      const size_t itf = (size_t)(item - d.item0) * d.nf + f;
      const double ah = att_hydro[out(z)], aa = att_atmo[out(z)];
      double pia = 0.0;
      if (s.radar_att == 1) {
        if (z < L - 1) pia = 2 * (0.0 + sa);
        pia = pia + ah + aa;
      } else if (s.radar_att == 2) {
        if (z > 0) pia = 2 * (sh + sa);
        pia = pia + ah + aa;
      }
      pia = pow(10.0, 0.1 * pia);
      for (int ip = 0; ip < np; ip++) {
        const double b = (s.radar_pol[ip] <= 2) ? w.back[itf] : 0.0;
        double zb = zconst * b * wl4;
        zb = zb / pia;
        ze[out(z) * np + ip] = (zb == 0.0) ? -9999.0 : 10 * log10(zb);
      }
    }
    if (td) {
      sa += att_atmo[out(z)];
    } else {
      sh += att_hydro[out(z)];
      sa += att_atmo[out(z)];
    }
  }
}

void launch_radar(const Dims &d, const Settings &s, const double *freqs, const Work &w,
                  double *ze, double *att_hydro, double *att_atmo, double *radar_hgt,
                  hipStream_t st) {
  const size_t n = (size_t)d.ncol_chunk * d.nf;
  if (n == 0 || !ze || !att_hydro || !att_atmo || !radar_hgt) return;
  hipLaunchKernelGGL(k_radar, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, d, s, freqs, w,
                     ze, att_hydro, att_atmo, radar_hgt);
}

}  // namespace pam
