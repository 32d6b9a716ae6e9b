/*
 * pamtra_amd.h -- C-ABI of the MI355X passive forward operator.
 *
 * This is the drop-in boundary for the work the reference does inside
 * `pypamtralib.run_pamtra()` (src/pyPamtraLib.f90:18-148) for passive-only
 * runs: run_rt (src/run_rt.f90:1-208) -> hydrometeor_extinction / Mie /
 * scatcnv -> rt4 / radtran4 -> collect_output.  The reference binds that
 * routine through f2py module globals which python/pyPamtra/libWrapper.py:24-235
 * fills from the `nmlSet`, descriptor-file and profile dictionaries.  Here the
 * same information travels as flat C-contiguous arrays with explicit shapes
 * (numpy's default C order of the pyPamtra dictionaries), strings are replaced
 * by enums, and nothing is held in global state.
 *
 * Plain pointers and sizes only: no torch or HIP types appear in the
 * signatures.  All arrays are C order (last index fastest).
 */
#ifndef PAMTRA_AMD_H
#define PAMTRA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Quadrature and polarisation of the RT4 solver (src/settings.f90:12-31).
 * The reference fixes nummu = 16 at compile time (settings.f90:16) and so
 * does libpamtra_amd.so; libpamtra_amd_n64.so is the same source built with
 * PAM_NUMMU 32 (32 angles per hemisphere, 64x64 operators), and a caller of
 * that library compiles this header with -DPAM_NUMMU=32. */
#ifndef PAM_NUMMU
#define PAM_NUMMU 16
#endif
#define PAM_NSTOKES 2
#define PAM_NANGLES (2 * PAM_NUMMU)
#define PAM_MAXNLEG 200
/* size bins per hydrometeor accepted by the device (the reference allocates
 * any nbin; its Mie path is bounded by the memory of one column) */
#define PAM_MAX_NBIN 4000

/* Error status codes (src/report_module.f90:9-14). */
enum pam_status { PAM_SUCCESS = 0, PAM_WARNING = 1, PAM_FATAL = 2, PAM_INFO = 3 };

/* descriptor file column "dist_name" (src/make_dist_param.f90). */
enum pam_dist_name {
  PAM_DIST_DISABLED = 0,
  PAM_DIST_MONO = 1,
  PAM_DIST_EXP = 2,
  PAM_DIST_EXP_FIELD_T = 3,
  PAM_DIST_EXP_RYAN = 4,
  PAM_DIST_MGAMMA = 5,
  PAM_DIST_LOGN = 6,
  PAM_DIST_EXP_COSMO_SNOW = 7,  /* COSMO two-moment snow, n_0(T, q) (make_dist_param.f90:343-374) */
  PAM_DIST_MONO_COSMO_ICE = 8,  /* COSMO ice, N(T), D <= 200 um (make_dist_param.f90:163-178) */
  PAM_DIST_MONO_ECHAM_CL = 9,   /* ECHAM6 liquid cloud, land/sea N (make_dist_param.f90:180-200) */
  PAM_DIST_MONO_ECHAM_ICE = 10, /* ECHAM6 ice volume diameter (make_dist_param.f90:202-209) */
  PAM_DIST_MGAMMA_MNH = 11,     /* Meso-NH, N_tot = C lambda^x (make_dist_param.f90:486-497) */
  PAM_DIST_NORM_GAMMA = 12,     /* normalised gamma, D_m and N_0* (make_dist_param.f90:537-600) */
  PAM_DIST_NORM_MGAMMA = 13     /* normalised modified gamma (make_dist_param.f90:602-624) */
};

/* descriptor file column "scat_name": "mie-sphere" (src/scatProperties.f90:722-769),
 * "ssrg-rt3[_kappa_beta_gamma[_zeta]]", the self-similar Rayleigh-Gans
 * aggregate model in its Legendre-series form (src/scatProperties.f90:771-863,
 * src/rayleigh_gans.f90:467-872), and "tmatrix", spheroids of axis ratio
 * as_ratio and canting angle canting by the extended-boundary-condition T
 * matrix (src/scatProperties.f90:289-368, src/tmatrix.f90:14-684), which adds
 * its scattering matrix, per-angle extinction matrix and emission vector to
 * the layer directly.  A call whose hydrometeors mix "tmatrix" with the
 * Legendre-series models is refused (PAM_FATAL): the reference's result
 * there reads an unset albedo (src/scatProperties.f90:141,911). */
enum pam_scat_name {
  PAM_SCAT_DISABLED = 0,
  PAM_SCAT_MIE_SPHERE = 1,
  PAM_SCAT_SSRG_RT3 = 2,
  PAM_SCAT_TMATRIX = 3
};

/* arithmetic of the device path */
enum pam_precision { PAM_PREC_FP64 = 0, PAM_PREC_MIXED = 1 };

/* nmlSet entries that reach the forward operator (python/pyPamtra/core.py:88-157,
 * defaults src/settings.f90:454-573).  gas_mod: 0 = 'R98'; liq_mod: 0 = 'TKC';
 * outpol: 0 = 'VH', 1 = 'IQ'.  active = 1 adds the radar simulator in its
 * "simple" mode (src/radar_simulator.f90:94-170): reflectivity Ze from the
 * Mie backscattering, hydrometeor and gas attenuation, radar heights.
 * radar_attenuation: 0 'disabled', 1 'top-down', 2 'bottom-up';
 * radar_pol[0..radar_npol-1]: 0 'NN', 1 'HH', 2 'VV', 3 'HV', 4 'VH';
 * radar_k2: |K|^2 of the radar equation per frequency (nfreq values, NULL:
 * 0.93 for every frequency, src/settings.f90:530). */
typedef struct pam_settings {
  int32_t passive;
  int32_t active;
  int32_t lgas_extinction;
  int32_t lhyd_absorption;
  int32_t lhyd_scattering;
  int32_t lhyd_emission;
  int32_t hydro_includehydroinrhoair;
  int32_t hydro_limit_density_area;
  int32_t hydro_adaptive_grid;
  int32_t conserve_mass_rescale_dsd;
  int32_t gas_mod;
  int32_t liq_mod;
  int32_t outpol;
  int32_t precision;
  double hydro_threshold;
  double hydro_softsphere_min_density;
  double emissivity;
  int32_t radar_attenuation;
  int32_t radar_npol;
  int32_t radar_pol[5];
  int32_t radar_pad_;
  const double *radar_k2;
} pam_settings;

/* One row of the 20-column descriptor file
 * (python/pyPamtra/descriptorFile.py:19-20).  hydro_name and vel_size_mod do
 * not enter the passive path and are not carried. */
typedef struct pam_hydro_desc {
  double as_ratio;
  int32_t liq_ice;
  int32_t moment_in;
  double rho_ms;
  double a_ms;
  double b_ms;
  double alpha_as;
  double beta_as;
  int32_t nbin;
  int32_t dist_name;
  double p_1, p_2, p_3, p_4;
  double d_1, d_2;
  int32_t scat_name;
  int32_t pad_;
  double canting;
  /* PAM_SCAT_SSRG_RT3 only: kappa, beta, gamma, zeta of the scattering
   * name (defaults 0.25, 1.0, 1.66 and 0.04 rounded to single precision,
   * src/scatProperties.f90:793-797) */
  double rg_kappa, rg_beta, rg_gamma, rg_zeta;
} pam_hydro_desc;

/* Optional per-(x,y,z,hydro) overrides of descriptor columns, the
 * descriptorFile4D of libWrapper.py:112-116 (hydrometeor_extinction.f90:153-245).
 * Each pointer is NULL or an array of shape (ngridx, ngridy, max_nlyrs, nhydro). */
typedef struct pam_hydro_4d {
  const double *as_ratio;
  const double *rho_ms;
  const double *a_ms;
  const double *b_ms;
  const double *alpha_as;
  const double *beta_as;
  const double *p_1, *p_2, *p_3, *p_4;
  const double *d_1, *d_2;
  const int32_t *nbin;
  /* Full-spectrum input (nmlSet hydro_fullspec with pyPamtra
   * df.addFullSpectra; src/hydrometeor_extinction.f90:120-150,
   * src/scatProperties.f90:258-270).  fs_nbin > 0 replaces the size
   * distribution of every hydrometeor by fs_nbin given bins: arrays of shape
   * (ngridx, ngridy, max_nlyrs, nhydro, fs_nbin), fs_d_bound with fs_nbin + 1
   * edges.  fs_n is the number per bin [m^-3] (not normalised); fs_rho and
   * fs_d the density and diameter the Mie spheres take; fs_mass, fs_as_ratio,
   * fs_canting and fs_rg_* the self-similar Rayleigh-Gans inputs. */
  int32_t fs_nbin;
  int32_t fs_pad_;
  const double *fs_rho, *fs_d, *fs_d_bound, *fs_n, *fs_mass, *fs_as_ratio, *fs_canting;
  const double *fs_rg_kappa, *fs_rg_beta, *fs_rg_gamma, *fs_rg_zeta;
} pam_hydro_4d;

/* The profile dictionary `pam.p` (python/pyPamtra/core.py:756-976).  NaN
 * marks a missing value that fillMissing_atmosphere_vars derives
 * (src/vars_atmosphere.f90:671-901).  Layer arrays may be NULL (all missing).
 * sfc_emissivity is an extension: a per-column emissivity for sfc_type=-9999
 * columns (NULL: settings.emissivity for every column, the reference's only
 * form, src/sfc_optics.f90:79-81). */
typedef struct pam_profile {
  int32_t ngridx, ngridy, max_nlyrs, noutlevels, nhydro;
  int32_t pad_;
  const int32_t *nlyrs;          /* (nx, ny) */
  const double *hgt_lev;         /* (nx, ny, L+1) [m] */
  const double *temp_lev;        /* (nx, ny, L+1) [K] */
  const double *press_lev;       /* (nx, ny, L+1) [Pa] */
  const double *relhum_lev;      /* (nx, ny, L+1) [%] */
  const double *hgt;             /* (nx, ny, L) or NULL */
  const double *temp;            /* (nx, ny, L) or NULL */
  const double *press;           /* (nx, ny, L) or NULL */
  const double *relhum;          /* (nx, ny, L) or NULL */
  const double *hydro_q;         /* (nx, ny, L, nhydro) [kg/kg] */
  const double *hydro_n;         /* (nx, ny, L, nhydro) [#/kg] */
  const double *hydro_reff;      /* (nx, ny, L, nhydro) [m] */
  const double *groundtemp;      /* (nx, ny) [K], NaN: temp_lev at the ground */
  const double *wind10u;         /* (nx, ny) [m/s] */
  const double *wind10v;         /* (nx, ny) [m/s] */
  const double *sfc_salinity;    /* (nx, ny) [ppt] */
  const double *sfc_emissivity;  /* (nx, ny) or NULL */
  const int32_t *sfc_type;       /* (nx, ny): -9999 constant emissivity, 0 ocean */
  const int32_t *sfc_model;      /* (nx, ny): -9999/0 TESSEM2 */
  const char *sfc_refl;          /* (nx, ny) one byte each: 'S', 'F' or 'L' */
  const double *obs_height;      /* (nx, ny, noutlevels) [m] */
} pam_profile;

/* Result arrays (vars_output.f90:134-141, keys of libWrapper.py:214-223).
 * Every pointer except tb may be NULL.  Diagnostic per-layer arrays are
 * indexed bottom-up like the profile. */
typedef struct pam_results {
  double *tb;            /* (nx, ny, noutlevels, 32, nfreq, 2) [K] */
  double *emissivity;    /* (nx, ny, 2, nfreq, 16) */
  double *angles_deg;    /* (32) */
  int32_t *status;       /* (nx, ny, nfreq) */
  double *kextatmo;      /* (nx, ny, L, nfreq) gas extinction [Np/m] */
  double *kexttot;       /* (nx, ny, L, nfreq) hydrometeor extinction [1/m] */
  double *salbedo;       /* (nx, ny, L, nfreq) single scattering albedo */
  int32_t *nlegen;       /* (nx, ny, L, nfreq) Legendre order, -1 clear */
  int32_t *ndoubles;     /* (nx, ny, L, nfreq) doublings, -1 clear layer */
  char *msg;             /* first error message, NUL terminated */
  int32_t msglen;
  int32_t pad_;
  /* radar outputs (settings.active = 1; vars_output.f90:144-155): -9999 in
   * Ze and radar_hgt where the reference leaves its fill value, NaN where a
   * (column, frequency) failed */
  double *ze;            /* (nx, ny, L, nfreq, radar_npol, 1) [dBz] */
  double *att_hydro;     /* (nx, ny, L, nfreq, 1) [dB] one way */
  double *att_atmo;      /* (nx, ny, L, nfreq) [dB] one way */
  double *radar_hgt;     /* (nx, ny, L) [m] */
} pam_results;

typedef struct pam_context pam_context;

/* Library version string. */
const char *pam_version(void);

/* PAM_NUMMU of the library (16 or 32): the angle count of its results.  No
 * counterpart in the reference, whose count is a compile-time PARAMETER
 * (settings.f90:16) that pyPamtra mirrors as _nangles (core.py:263). */
int32_t pam_nummu(void);

/* Number of HIP devices visible to the library (0 when there is none). */
int pam_device_count(void);

/* Creates a context bound to one HIP device.  Replaces the module-global
 * state the reference resets in deallocate_everything (libWrapper.py:48).
 * Returns PAM_SUCCESS or PAM_FATAL (no device, allocation failure). */
int pam_context_create(pam_context **ctx, int device);

/* Frees the device workspace and the context. */
int pam_context_destroy(pam_context *ctx);

/* Runs the passive forward operator for every (column, frequency):
 * the replacement of run_pamtra (src/pyPamtraLib.f90:119-133), host arrays in
 * and out.  freqs in GHz, sorted ascending (core.py:1536).  Returns the worst
 * status over all columns: PAM_SUCCESS, PAM_WARNING or PAM_FATAL; never
 * aborts the process (radmat.f90:180-189 STOPs, this returns PAM_FATAL).
 * The work runs on the context's own non-blocking stream: different
 * contexts may run concurrently, one host thread per context. */
int pam_run(pam_context *ctx, const pam_settings *settings,
            const pam_hydro_desc *hydros, const pam_hydro_4d *hydro4d,
            const pam_profile *profile, const double *freqs, int32_t nfreq,
            pam_results *results);

/* As pam_run, but every array in `profile`, `hydro4d` and `results` (except
 * msg) is device memory already resident on the context's device, and the
 * work is queued on `stream` (a hipStream_t, NULL for the null stream).
 * freqs and hydros stay host arrays.  The call returns after the results are
 * written (it synchronises `stream`). */
int pam_run_device(pam_context *ctx, const pam_settings *settings,
                   const pam_hydro_desc *hydros, const pam_hydro_4d *hydro4d,
                   const pam_profile *profile, const double *freqs,
                   int32_t nfreq, pam_results *results, void *stream);

/* As pam_run_device, but allocating nothing: no device memory, no events,
 * no graphs.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Its workspace is a fixed array of
 * PAM_NOALLOC_SCRATCH_BYTES per device, loaded with the library and held
 * until the process ends; columns are processed in chunks that fit it, and
 * a call whose smallest chunk (one column's optics) does not fit is refused
 * with PAM_FATAL and a message.  Every call on one device shares that array
 * and leaves its contents behind: calls that overlap on one device (other
 * host threads or contexts, or calls on different streams) race.  The call
 * synchronises `stream` before it returns; the stage times of
 * pam_last_stage_times are not recorded. */
#define PAM_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int pam_run_device_noalloc(pam_context *ctx, const pam_settings *settings,
                           const pam_hydro_desc *hydros, const pam_hydro_4d *hydro4d,
                           const pam_profile *profile, const double *freqs,
                           int32_t nfreq, pam_results *results, void *stream);

/* Kernel-level timings of the last call (milliseconds, measured with HIP
 * events on the call's stream), in the order of pam_stage_names(). */
/* calc_tmatrix (src/tmatrix.f90:14-444) for one hydrometeor on the device:
 * the sums over nbins size bins of the T-matrix spheroid's scattering
 * matrix, extinction matrix and emission vector times ndens * del_d, in the
 * reference's Fortran order, smat (2,NUMMU,2,NUMMU,2), ext (2,2,NUMMU),
 * emis (2,NUMMU), NUMMU = pam_nummu().  freq_hz in Hz; (mre, mim) the
 * refractive index m = mre - i mim of scatProperties.f90:243; phase
 * (liq_ice) -1 mixes ice of density below 917 kg/m3 with air; dmax, del_d
 * [m], ndens [m^-4], density [kg/m3], as_ratio and canting [deg] per bin.
 * Host arrays.  nmax (optional,
 * nbins entries) receives each bin's T-matrix order.  Returns PAM_SUCCESS
 * or PAM_FATAL (invalid bins, a particle beyond the device's limits: order
 * 40, 100 Gauss points). */
int pam_calc_tmatrix(pam_context *ctx, double freq_hz, double mre, double mim, int32_t phase,
                     int32_t nbins, const double *dmax, const double *del_d,
                     const double *ndens, const double *density, const double *as_ratio,
                     const double *canting, double *smat, double *ext, double *emis,
                     int32_t *nmax);

int pam_last_stage_times(const pam_context *ctx, double *ms, int32_t n);
const char *pam_stage_names(void);

/* Work counters of the last call: [0] columns, [1] channels,
 * [2] scattering (layer, channel) items, [3] total doubling steps,
 * [4] scattering adding steps, [5] clear adding steps, [6] Mie bins,
 * [7] sum over Mie bins of the series length, [8] sum over scattering items
 * of (NUMPTS/2+1)*(nlegen+1), the phase-matrix work, [9] 32x32 matrix
 * products executed by the doubling, [10] the same for the adding,
 * [11] units recomputed by the Gauss-Jordan pass, [12] sum over Mie bins of
 * nquad * series length (the angular sums), [13] sum over (item,
 * hydrometeor) of nquad * (nlegen + 1) (the Legendre projection). */
int pam_last_work(const pam_context *ctx, int64_t *counts, int32_t n);

/* Device bytes one call needs for these dimensions, with `nitems` layers
 * holding hydrometeors above hydro_threshold (an upper bound if nitems is).
 * Used for admission by an HBM budget.  No counterpart in the reference,
 * which runs one call per process (python/pyPamtra/core.py:1613-1647). */
int64_t pam_workspace_bytes(int64_t ncol, int32_t max_nlyrs, int32_t nhydro,
                            int32_t nbin_max, int32_t nfreq, int32_t noutlevels,
                            int64_t nitems);

/* Caps the per-layer optics workspace of each call on this context (the
 * columns are then processed in chunks that fit).  Default: half of the free
 * HBM at creation, at most 96 GiB; PAM_WORKSPACE_GB overrides both.  No
 * counterpart in the reference. */
int pam_context_set_workspace_limit(pam_context *ctx, int64_t bytes);

/* Device bytes the context's grow-only workspace holds now.  No counterpart
 * in the reference. */
int64_t pam_context_bytes(const pam_context *ctx);

/* Releases the context's grow-only workspace (the device's Mie tables stay).
 * No counterpart in the reference (its module globals are reallocated per
 * call, libWrapper.py:48). */
int pam_context_trim(pam_context *ctx);

/* Launch shapes of the context, chosen by pamtra_amd.autotune or set by the
 * caller; results are bitwise the same for every value.  Keys
 * (pam_tuning_keys, comma-separated):
 *   mie_persistent, doubling_persistent, adding_persistent: 0 launches one
 *     workgroup per unit; k in 1..64 launches k workgroups per compute unit
 *     that take units from a work queue;
 *   graph: 1 captures the kernels of a single-chunk call in a HIP graph and
 *     replays it while the launch arguments are unchanged;
 *   pipeline: k in 2..16 (the default 0: off) splits a call of at least 8192 k
 *     units into at least k column chunks run on two streams (the context's
 *     second stream beside the call's), each chunk's DSD and Mie beside the
 *     previous chunk's phase matrix, doubling and adding; 0 or 1 runs the
 *     call's chunks in order on its one stream.  The stage times of a
 *     pipelined call are the sums of each stream's segments and overlap.
 * pam_context_set_tuning returns PAM_FATAL for an unknown key or value;
 * pam_context_get_tuning returns -1 for an unknown key.  No counterpart in
 * the reference (its Fortran loops have no launch shape). */
const char *pam_tuning_keys(void);
int pam_context_set_tuning(pam_context *ctx, const char *key, int64_t value);
int64_t pam_context_get_tuning(const pam_context *ctx, const char *key);

#ifdef __cplusplus
}
#endif

#endif /* PAMTRA_AMD_H */
