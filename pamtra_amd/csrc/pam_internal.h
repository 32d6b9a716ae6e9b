// pam_internal.h -- device workspace layout, launch helpers and device math
// shared by the HIP translation units of libpamtra_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/pamtra_amd.h"

namespace pam {

constexpr int NMU = PAM_NUMMU;       // quadrature angles per hemisphere
constexpr int NST = PAM_NSTOKES;     // Stokes parameters
constexpr int NN = NMU * NST;        // order of the R / T matrices (32)
constexpr int NN2 = NN * NN;
// global scratch of the adding kernels at N = 64, per workgroup slot (one
// slot per CU): RT_GMATS padded NN x (NN + 1) FP64 matrices (none at N = 32)
constexpr int RT_GMATS = (NN > 32) ? 6 : 0;
constexpr size_t RT_SLOT_DOUBLES = (size_t)RT_GMATS * NN * (NN + 1);
constexpr int MAXNLEG = PAM_MAXNLEG; // Legendre coefficients per series
constexpr int NCOEF = 3;             // P11, P12, P33 series
constexpr int MAX_NQUAD = 256;       // Gauss nodes of the Mie phase integral
constexpr int MAX_NTERMS = 160;      // Mie series length supported on device
constexpr int MAX_HYDRO = 32;        // bits of the per-layer hydrometeor mask
constexpr int MAX_NUMPTS = 204;      // 2*int((199+1)/2)+4
constexpr int PH_NBUCKET = MAX_NUMPTS / 2 - 1;  // NUMPTS values 4, 6, ..., 204
// T-matrix producer: largest order NMAX and number of Gauss points NGAUSS
// (the reference allows 99 and 500); a particle that needs more is refused
constexpr int TM_NCAP = 40;
constexpr int TM_GCAP = 100;
constexpr int TM_NSMAT = 2 * NMU * 2 * NMU * 2;  // scattering matrix (2,NMU,2,NMU,2)
constexpr int TM_PER_CU = 8;  // one-wave workgroups per CU of its persistent grid

constexpr double PI = 3.141592653589793;      // constants.f90:16
constexpr double C_LIGHT = 299792458.0;       // constants.f90:15
constexpr double R_D = 287.0596736665907;
constexpr double R_V = 461.5249933083879;
constexpr double SKY_TEMP = 2.73;
constexpr double DELTA_D_MONO = 1.0e-8;

// error codes carried in the status arrays (high bits) next to the
// reference's severity (low byte)
enum ErrCode : int32_t {
  E_NONE = 0,
  E_NLYRS = 1 << 8,
  E_NAN_LEVEL = 2 << 8,
  E_NAN_LAYER = 3 << 8,
  E_HGT_ORDER = 4 << 8,
  E_GAS = 5 << 8,
  E_SURFACE = 6 << 8,
  E_MASS_SIZE = 7 << 8,
  E_DIST_PARAMS = 8 << 8,
  E_ADAPTIVE = 9 << 8,
  E_MOMENT = 10 << 8,
  E_SOFT = 11 << 8,
  E_MIE_INPUT = 12 << 8,
  E_MIE_TERMS = 13 << 8,
  E_MIE_RESULT = 14 << 8,
  E_KEXT = 15 << 8,
  E_NLEG = 16 << 8,
  E_NORM = 17 << 8,
  E_SINGULAR = 18 << 8,
  E_REFL = 19 << 8,
  E_NAN_COEF = 20 << 8,
  E_SSRG_INPUT = 21 << 8,
  E_TMATRIX = 22 << 8,
};

enum FaultKind : int { FAULT_NONE = 0, FAULT_NAN_COEF = 1, FAULT_SINGULAR = 2 };

// one call's shapes and settings, passed by value to every kernel
struct Dims {
  int nx, ny, ncol, ML, nout, nh, nf, nbmax;
  int col0;         // first column of the current chunk (global index)
  int ncol_chunk;   // columns in the chunk
  int nitems;       // scattering layers in the chunk
  int item0;        // global index of the chunk's first item
};

struct Settings {
  int lgas, lhyd_abs, lhyd_scat, lhyd_emis, incl_hydro_rho, limit_density, adaptive,
      conserve, outpol, precision;
  // inverses: 0 series with a Gauss-Jordan pass for the units that do not
  // converge; 1 Gauss-Jordan only (PAM_INVERSE=gauss-jordan); 2 every unit
  // through the Gauss-Jordan queue (PAM_INVERSE=queue-all, a test mode)
  int inverse_gj;
  // launch shape: workgroups per compute unit of a persistent grid that
  // takes units from a work queue (0: one workgroup per unit)
  int pers_mie, pers_dbl, pers_add;
  int ncu;  // compute units of the device
  int ssrg;  // some hydrometeor uses the self-similar Rayleigh-Gans model
  // 1: the chunked Mie kernel for every call (PAM_MIE=chunked, a test mode);
  // otherwise calls without self-similar Rayleigh-Gans take the
  // coefficient-first kernel
  int mie_chunked;
  // some hydrometeor uses the T-matrix producer (then every scattering one
  // does: calls mixing it with the Legendre-series models are refused); 2
  // when one of them has a canting angle
  int tmatrix;
  // 1: the mixed-precision path keeps the FP64 Mie kernel (PAM_MIE=fp64)
  int mie_fp64;
  // test mode (PAM_TEST_FAULT=<kind>:<column>:<channel>): the named (column,
  // channel) meets the failure the reference stops on, so that its isolation
  // can be tested; kind FAULT_NAN_COEF makes its Legendre coefficients NaN
  // (scatcnv.f90:38-43), FAULT_SINGULAR its surface inverse singular
  // (radmat.f90:180-189)
  int fault, fault_col, fault_ch;
  // radar simulator, "simple" mode (active): attenuation 0 disabled,
  // 1 top-down, 2 bottom-up; polarisations 0 NN, 1 HH, 2 VV, 3 HV, 4 VH
  int active, radar_att, radar_npol, radar_pol[5];
  double threshold, softsphere_min, emissivity;
};

// input profile (device pointers), see pam_profile
struct Profile {
  const int32_t *nlyrs;
  const double *hgt_lev, *temp_lev, *press_lev, *relhum_lev;
  const double *hgt, *temp, *press, *relhum;
  const double *hydro_q, *hydro_n, *hydro_reff;
  const double *groundtemp, *wind10u, *wind10v, *sfc_salinity, *sfc_emissivity;
  const int32_t *sfc_type, *sfc_model;
  const char *sfc_refl;
  const double *obs_height;
};

struct Hydro4D {
  const double *as_ratio, *rho_ms, *a_ms, *b_ms, *alpha_as, *beta_as, *p_1, *p_2, *p_3, *p_4,
      *d_1, *d_2;
  const int32_t *nbin;
  // full spectrum given per (column, layer, hydrometeor, bin); fs_nbin 0: none
  int fs_nbin;
  const double *fs_rho, *fs_d, *fs_d_bound, *fs_n, *fs_mass, *fs_as_ratio, *fs_canting,
      *fs_rg_kappa, *fs_rg_beta, *fs_rg_gamma, *fs_rg_zeta;
};

// device workspace (all device pointers)
struct Work {
  // per column (global column index)
  double *hgt_lev, *temp_lev;          // [ncol][ML+1]
  double *temp, *press, *rho_v, *q_hum, *dz;  // [ncol][ML]
  double *groundtemp;                  // [ncol]
  double *hgt;                         // [ncol][ML] layer heights
  int32_t *nl;                         // [ncol]
  int32_t *outlev;                     // [ncol][nout]
  int32_t *col_status;                 // [ncol]
  uint32_t *hmask;                     // [ncol][ML]
  int32_t *col_nitems;                 // [ncol]
  int32_t *col_item0;                  // [ncol] exclusive prefix
  int32_t *layer_item;                 // [ncol][ML], -1 clear
  // per column and frequency
  double *kgas;                        // [ncol][nf][ML]
  double *sfc_emis, *sfc_refl;         // [ncol][nf][2][NMU]
  int32_t *cf_status;                  // [ncol][nf]
  // per item of the chunk
  int32_t *item_col, *item_layer;      // [nitems]
  // dsd_rho: particle density for Mie, particle mass for self-similar
  // Rayleigh-Gans hydrometeors
  double *dsd_d, *dsd_dd, *dsd_nd, *dsd_rho;  // [nitems][nh][nbmax]
  int32_t *dsd_flag;                   // [nitems][nh]
  double *dsd_ar;                      // [nitems][nh] as_ratio of the layer
  // per (item, frequency) of the chunk
  double *kext, *salb;                 // [nitems*nf]
  double *back;                        // [nitems*nf] co-polar backscattering [m2/m3]
  int32_t *nleg;                       // [nitems*nf]
  double *coef;                        // [nitems*nf][3][MAXNLEG]
  int32_t *ndbl;                       // [nitems*nf]
  int32_t *itf_status;                 // [nitems*nf]
  double *R, *T;                       // [nitems*nf][NN2]
  double *Sp, *Sm;                     // [nitems*nf][NN]
  // Mie angle tables per series length N (see build_angle_tables)
  const double *atab;
  const float *atab_f;                 // the same tables rounded to FP32
  double *mie_m;                       // per unit: water, ice refractive index (re, im)
  const int64_t *atab_off;             // [atab_nmax + 1]
  int atab_nmax;
  // phase-matrix tables per NUMPTS (see build_phase_tables) and the units
  // of the chunk gathered by NUMPTS
  const double *ptab;
  const int64_t *ptab_off;             // [PH_NBUCKET + 1]
  const float *ptab_f;                 // the same tables rounded to FP32 (mixed path)
  uint32_t *ph_cnt, *ph_off, *ph_toff; // [PH_NBUCKET + 1]
  int32_t *ph_rank, *ph_perm;          // [nitems*nf]
  const double *mu, *qw;               // Lobatto nodes/weights [NMU]
  const double *radar_k2;              // [nf] |K|^2 of the radar equation
  // units whose inverse series did not converge, for the Gauss-Jordan pass
  uint32_t *redo_n;                    // [3]: doubling, adding, levels (reset per chunk)
  int32_t *redo_dbl;                   // [nitems*nf] of the chunk
  int32_t *redo_add;                   // [ncol*nf]
  int32_t *redo_lev;                   // [ncol*nf]: units with an interior output level
  double *rt_scratch;                  // [ncu][RT_SLOT_DOUBLES] at N = 64, else null
  // U slots of the coefficient-first Mie kernel: one per workgroup of its
  // persistent grid, mie_ld rows (16 per 8 bins) x 2 MAX_NTERMS columns
  double *mie_u;
  size_t mie_slot_doubles;
  int mie_ld, mie_slots;
  // T-matrix producer: per-angle extinction (k11, k12) and emission (e1, e2)
  // of each unit [nitems*nf][NMU][2], null when no hydrometeor uses it; the
  // producer's workgroup slots, the call's geometry table and the Gauss
  // node table (see pam_tmatrix.hip)
  double *tm_ext, *tm_emis;
  double *tm_scr;
  int tm_slots;
  double *tm_geo;
  const double *tm_gauss;
  // work counters (atomics)
  unsigned long long *counters;        // [PAM_NWORK], see pam_last_work
  // work-queue counters of the persistent grids: mie, doubling, adding
  // (cleared per chunk)
  uint32_t *queue;                     // [4]
};

// ------------------------------------------------------------------
// device math helpers
// ------------------------------------------------------------------

__device__ __forceinline__ double sp(double x) { return (double)(float)x; }
// The phase matrix k_phase leaves in a unit's R and T slots: FP32 in the
// mixed path (formed from the FP32 copy of its tables), FP64 otherwise.
__device__ __forceinline__ bool phase_in_f32(const Settings &s, const Work &w) {
  return s.precision == PAM_PREC_MIXED && w.ptab_f != nullptr;
}
// entry e of the quadrant stored at `slot`
__device__ __forceinline__ double phase_at(bool f32, const double *slot, int e) {
  return f32 ? (double)reinterpret_cast<const float *>(slot)[e] : slot[e];
}

struct cd {
  double re, im;
};
__device__ __forceinline__ cd mk(double r, double i) { return cd{r, i}; }
__device__ __forceinline__ cd operator+(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd operator-(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd operator*(cd a, cd b) {
  return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cd operator*(double s, cd a) { return cd{s * a.re, s * a.im}; }
__device__ __forceinline__ cd operator*(cd a, double s) { return cd{a.re * s, a.im * s}; }
__device__ __forceinline__ cd operator+(cd a, double s) { return cd{a.re + s, a.im}; }
__device__ __forceinline__ cd operator-(cd a, double s) { return cd{a.re - s, a.im}; }
__device__ __forceinline__ cd conj(cd a) { return cd{a.re, -a.im}; }
// Smith's algorithm for complex division
__device__ __forceinline__ cd cdiv(cd a, cd b) {
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, den = b.re + b.im * r;
    return cd{(a.re + a.im * r) / den, (a.im - a.re * r) / den};
  } else {
    double r = b.re / b.im, den = b.re * r + b.im;
    return cd{(a.re * r + a.im) / den, (a.im * r - a.re) / den};
  }
}
__device__ __forceinline__ double cabs_(cd a) { return hypot(a.re, a.im); }
__device__ __forceinline__ cd csqrt_(cd z) {
  // principal square root
  double m = hypot(z.re, z.im);
  double r = sqrt(0.5 * (m + fabs(z.re)));
  if (r == 0.0) return cd{0.0, 0.0};
  if (z.re >= 0.0) return cd{r, z.im / (2.0 * r)};
  return cd{fabs(z.im) / (2.0 * r), z.im >= 0.0 ? r : -r};
}

// Sihvola mixing with air applied to a refractive index (eps_mix.f90:1-48,
// called as in mie_spheres.f90:187-192); single-precision result
__device__ __forceinline__ cd eps_mix_air(cd e1, double rho) {
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

// Planck radiance [W m^-2 sr^-1 um^-1] (rt_utilities.f90:69-90, units 'r')
__device__ __forceinline__ double planck(double temp, double wl) {
  if (temp > 0.0) {
    double w2 = wl * wl, w5 = w2 * w2 * wl;
    return 1.1911e8 / w5 / (exp(1.4388e4 / (wl * temp)) - 1.0);
  }
  return 0.0;
}

// inverse Planck (CONVERT_OUTPUT, rt_utilities.f90:705-735)
__device__ __forceinline__ double rad2tb(double rad, double wl) {
  double w2 = wl * wl, w5 = w2 * w2 * wl;
  if (rad > 0.0) return 1.4388e4 / (wl * log(1.0 + 1.1911e8 / (rad * w5)));
  if (rad == 0.0) return 0.0;
  return -1.4388e4 / (wl * log(1.0 + 1.1911e8 / (-rad * w5)));
}

__device__ __forceinline__ void set_err(int32_t *st, int32_t code) {
  atomicMax(st, code | PAM_FATAL);
}

// Next unit of a persistent workgroup: thread 0 takes an index from the
// queue counter and the workgroup reads it from LDS (the first barrier ends
// the previous unit's use of LDS).  Every workgroup leaves its loop once the
// counter has passed the unit count.
__device__ __forceinline__ int next_unit(uint32_t *ctr, int *slot) {
  __syncthreads();
  if (threadIdx.x == 0) *slot = (int)atomicAdd(ctr, 1u);
  __syncthreads();
  return *slot;
}

// Grid of a launch over n units: one workgroup per unit, or per_cu
// workgroups per compute unit (at most n) for a persistent grid.
inline size_t persistent_grid(size_t n, int per_cu, int ncu) {
  if (per_cu <= 0 || ncu <= 0) return n;
  const size_t g = (size_t)per_cu * (size_t)ncu;
  return g < n ? g : n;
}

}  // namespace pam

// ------------------------------------------------------------------
// launchers (host side), one per translation unit
// ------------------------------------------------------------------
namespace pam {
void launch_columns(const Dims &d, const Settings &s, const Profile &p, const pam_hydro_desc *hyd,
                    const Hydro4D &h4, const double *freqs, Work &w, hipStream_t st);
void launch_items(const Dims &d, Work &w, hipStream_t st);
void launch_dsd(const Dims &d, const Settings &s, const Profile &p, const pam_hydro_desc *hyd,
                const Hydro4D &h4, Work &w, hipStream_t st);
// Mie angle tables: per series length N = 1..nmax the Gauss-Legendre nodes
// and weights of the phase integral, pi_n/tau_n at the nodes, and the
// Legendre polynomials at the nodes.  angle_table_layout fills off[0..nmax]
// (doubles) and returns the total size; build_angle_tables fills the table.
int64_t angle_table_layout(int64_t *off, int *nmax);
void build_angle_tables(double *tab, const int64_t *d_off, int nmax, hipStream_t st);
// dst[i] = (float)src[i], i < n: the FP32 copy of the angle tables
void round_tables_f32(const double *tab, float *tab_f, int64_t n, hipStream_t st);
// phase-matrix tables: off[0..PH_NBUCKET] (doubles), total size returned;
// d_mu holds the NMU Lobatto nodes
int64_t phase_table_layout(int64_t *off);
void build_phase_tables(double *tab, const int64_t *d_off, const double *d_mu, hipStream_t st);
// phase matrices of the chunk's scattering units into their R/T slots
void launch_phase(const Dims &d, const Settings &s, Work &w, hipStream_t st);
void launch_mie_index(const Dims &d, const pam_hydro_desc *hyd, const double *freqs, Work &w,
                      hipStream_t st);
void launch_tmatrix(const Dims &d, const Settings &s, const pam_hydro_desc *hyd,
                    const Hydro4D &h4, const double *freqs, Work &w, hipStream_t st);
size_t tm_slot_doubles();
size_t tm_geo_doubles();
std::vector<double> tm_gauss_table();
size_t tm_nacc();
void launch_tm_hydro(int nbins, const double *dmax, const double *del_d, const double *ndens,
                     const double *rho, const double *as_ratio, const double *canting,
                     double freq_hz, double mre, double mim, int phase, const double *mu,
                     const double *qw, const double *gtab, double *geo, double *scr, double *out,
                     int *status, int *nmax_out, hipStream_t st);
void launch_mie(const Dims &d, const Settings &s, const pam_hydro_desc *hyd, const Hydro4D &h4,
                const Profile &p, const double *freqs, Work &w, hipStream_t st);
// rows and doubles of one U slot of the coefficient-first Mie kernel
int mie_slot_rows(int nbmax);
size_t mie_slot_doubles(int nbmax);
constexpr int MIE2_PER_CU = 8;  // default workgroups per CU of its persistent grid
void launch_surface(const Dims &d, const Settings &s, const Profile &p, const double *freqs,
                    Work &w, double *emis_out, hipStream_t st);
void launch_doubling(const Dims &d, const Settings &s, const double *freqs, Work &w,
                     hipStream_t st);
void launch_adding(const Dims &d, const Settings &s, const double *freqs, const char *sfc_refl,
                   Work &w, double *tb, hipStream_t st);
// radar "simple" mode of the chunk's columns (after the adding, which
// settles each (column, channel)'s status)
void launch_radar(const Dims &d, const Settings &s, const double *freqs, const Work &w,
                  double *ze, double *att_hydro, double *att_atmo, double *radar_hgt,
                  hipStream_t st);
// n 32-bit words at p set to v (a kernel, so that it can sit in a captured
// graph)
void launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t st);
}  // namespace pam
