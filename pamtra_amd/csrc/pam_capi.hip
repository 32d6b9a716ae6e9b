// pam_capi.hip -- the C-ABI of include/pamtra_amd.h and the host-side
// scheduling of one call: the replacement of run_pamtra's frequency/column
// loop nest (pyPamtraLib.f90:119-133).  Columns are processed in chunks
// sized so that the per-layer optics of a chunk fit the device workspace
// budget; within a chunk every stage is one kernel launch over all
// (layer, channel) or (column, channel) units.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>


#include "pam_internal.h"

using namespace pam;

#define HIP_OK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      set_msg(res, std::string("HIP error: ") + hipGetErrorString(e_) + " at " \
                       #x);                                                    \
      return PAM_FATAL;                                                        \
    }                                                                          \
  } while (0)

constexpr int PAM_NWORK = 14;
static const char *k_stage_names = "columns,dsd,mie,phase,doubling,adding";

// The workspace of pam_run_device_noalloc: one array per device, loaded with
// the code object and held until the process ends.  Host code never passes
// `scratch` itself to a kernel (on the host it only names the symbol): its
// device address comes from hipGetSymbolAddress at pam_context_create, once
// per device, and is kept in g_scratch_ptr.
static __device__ __align__(256) unsigned char scratch[PAM_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mutex;
static std::map<int, void *> g_scratch_ptr;

// Device address of this device's `scratch` (the device is current); null on
// failure.
static void *scratch_address(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) return it->second;
  void *ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
    (void)hipGetLastError();
    return nullptr;
  }
  g_scratch_ptr[device] = ptr;
  return ptr;
}

// Workspace regions of one call: grown DevBufs of the context, or (arena
// set) consecutive pieces of a fixed region, with nothing allocated.
struct Arena {
  char *base = nullptr;
  size_t size = 0, used = 0;
  // n bytes at a 256-byte boundary, or null when they do not fit
  void *take(size_t n) {
    const size_t a = (used + 255) & ~(size_t)255;
    if (a > size || n > size - a) return nullptr;
    used = a + n;
    return base + a;
  }
  size_t left() const {
    const size_t a = (used + 255) & ~(size_t)255;
    return a >= size ? 0 : size - a;
  }
};
constexpr int NSTAGE = 6;
constexpr int PAM_NTUNE = 5;  // entries of k_tuning_keys
// pipelined calls (tuning key "pipeline" = k >= 2): at least k column chunks
// of at least PIPE_MIN_UNITS (column, channel) or (layer, channel) units
constexpr int PIPE_MAX = 16, PIPE_DEFAULT = 0;
constexpr int64_t PIPE_MIN_UNITS = 8192;

// A device allocation that only grows.
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t n) {
    if (n <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = n + n / 8 + 4096;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Mie angle tables of one device, shared by all its contexts and kept until
// the process ends (66 MiB, and their 33 MiB FP32 copy).
struct AngleTables {
  DevBuf buf, pbuf, fbuf, pfbuf, gbuf;
  const double *tm_gauss = nullptr;  // Gauss nodes of the T-matrix producer
  const double *tab = nullptr;
  const float *tab_f = nullptr;
  const int64_t *off = nullptr;
  int nmax = 0;
  const double *ptab = nullptr;  // phase-matrix tables (about 200 MiB)
  const float *ptab_f = nullptr;  // their FP32 copy (the mixed path)
  const int64_t *poff = nullptr;
};
static void lobatto(int num, double *mu, double *w);
static std::mutex g_tab_mutex;
static std::map<int, AngleTables *> g_tables;

// Builds the tables of the current device once; nullptr on failure.
static AngleTables *device_tables(int device) {
  std::lock_guard<std::mutex> lock(g_tab_mutex);
  auto it = g_tables.find(device);
  if (it != g_tables.end()) return it->second;
  AngleTables *t = new AngleTables();
  int64_t off[MAX_NTERMS + 1];
  int nmax = 0;
  const int64_t n = angle_table_layout(off, &nmax);
  const size_t tab_bytes = ((size_t)n * 8 + 255) & ~(size_t)255;
  bool ok = t->buf.ensure(tab_bytes + sizeof(off)) == hipSuccess;
  if (ok) {
    double *tab = (double *)t->buf.p;
    int64_t *doff = (int64_t *)((char *)t->buf.p + tab_bytes);
    ok = hipMemcpy(doff, off, sizeof(int64_t) * (nmax + 1), hipMemcpyHostToDevice) == hipSuccess;
    if (ok) ok = t->fbuf.ensure((size_t)n * 4) == hipSuccess;
    if (ok) {
      build_angle_tables(tab, doff, nmax, nullptr);
      round_tables_f32(tab, (float *)t->fbuf.p, n, nullptr);
      ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess;
    }
    t->tab = tab;
    t->tab_f = (const float *)t->fbuf.p;
    t->off = doff;
    t->nmax = nmax;
  }
  if (ok) {
    int64_t poff[PH_NBUCKET + 1];
    const int64_t pn = phase_table_layout(poff);
    const size_t pbytes = ((size_t)pn * 8 + 255) & ~(size_t)255;
    ok = t->pbuf.ensure(pbytes + sizeof(poff) + 256 + NMU * sizeof(double)) == hipSuccess;
    if (ok) {
      double *ptab = (double *)t->pbuf.p;
      int64_t *dpoff = (int64_t *)((char *)t->pbuf.p + pbytes);
      double *dmu = (double *)((char *)dpoff + ((sizeof(poff) + 255) & ~(size_t)255));
      double mu[NMU], qw[NMU];
      lobatto(NMU, mu, qw);
      ok = hipMemcpy(dpoff, poff, sizeof(poff), hipMemcpyHostToDevice) == hipSuccess &&
           hipMemcpy(dmu, mu, sizeof(mu), hipMemcpyHostToDevice) == hipSuccess;
      if (ok) ok = t->pfbuf.ensure((size_t)pn * 4) == hipSuccess;
      if (ok) {
        build_phase_tables(ptab, dpoff, dmu, nullptr);
        round_tables_f32(ptab, (float *)t->pfbuf.p, pn, nullptr);
        ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess;
      }
      t->ptab = ptab;
      t->ptab_f = (const float *)t->pfbuf.p;
      t->poff = dpoff;
    }
  }
  if (ok) {
    const std::vector<double> g = tm_gauss_table();
    ok = t->gbuf.ensure(g.size() * 8) == hipSuccess &&
         hipMemcpy(t->gbuf.p, g.data(), g.size() * 8, hipMemcpyHostToDevice) == hipSuccess;
    t->tm_gauss = (const double *)t->gbuf.p;
  }
  if (!ok) {
    t->buf.release();
    t->pbuf.release();
    t->fbuf.release();
    t->pfbuf.release();
    t->gbuf.release();
    delete t;
    return nullptr;
  }
  g_tables[device] = t;
  return t;
}

struct pam_context {
  int device = 0;
  hipStream_t stream = nullptr;  // the stream of pam_run (non-blocking)
  hipStream_t stream2 = nullptr; // the second stream of a pipelined call
  DevBuf colbuf, itembuf, iobuf, smallbuf;
  DevBuf rtbuf;  // global scratch of the adding kernels at N = 64
  DevBuf miebuf; // U slots of the coefficient-first Mie kernel
  DevBuf tmbuf;  // workgroup slots and geometry of the T-matrix producer
  const double *d_tm_gauss = nullptr;
  const double *d_atab = nullptr;
  const float *d_atab_f = nullptr;
  const int64_t *d_atab_off = nullptr;
  int atab_nmax = 0;
  const double *d_ptab = nullptr;
  const float *d_ptab_f = nullptr;
  const int64_t *d_ptab_off = nullptr;
  double *d_mu = nullptr, *d_qw = nullptr;
  double mu[NMU], qw[NMU];
  double stage_ms[NSTAGE] = {0, 0, 0, 0, 0, 0};
  int64_t work[PAM_NWORK] = {};
  size_t budget_bytes = 0;
  int ncu = 0;                     // compute units of the device
  int64_t tune[PAM_NTUNE] = {};    // launch shapes, see pam_context_set_tuning
  // graph of the last single-chunk call and the launch arguments it was
  // captured with (replayed while they are unchanged)
  hipGraphExec_t gexec[5] = {};   // one graph per timed stage of the chunk
  std::vector<unsigned char> gkey;
  void *scratch = nullptr;        // device address of this device's `scratch`
};

static const char *k_tuning_keys =
    "mie_persistent,doubling_persistent,adding_persistent,graph,pipeline";

static int tuning_index(const char *key) {
  if (!key) return -1;
  int i = 0;
  const char *p = k_tuning_keys;
  while (*p) {
    const char *q = strchr(p, ',');
    const size_t len = q ? (size_t)(q - p) : strlen(p);
    if (strlen(key) == len && strncmp(key, p, len) == 0) return i;
    if (!q) break;
    p = q + 1;
    i++;
  }
  return -1;
}

static void drop_graph(pam_context *c) {
  for (auto &g : c->gexec) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
  c->gkey.clear();
}

static void set_msg(pam_results *res, const std::string &m) {
  if (res && res->msg && res->msglen > 0) {
    snprintf(res->msg, (size_t)res->msglen, "%s", m.c_str());
  }
}

// Lobatto half-space quadrature (rt_utilities.f90:165-203), host side
static void lobatto(int num, double *mu, double *w) {
  int n = 2 * num, n1 = n - 1;
  double ci = (n % 2 == 1) ? 1.0 : 0.5;
  for (int j = 1; j <= num - 1; j++) {
    double x = sin((double)3.141592654f * (j - ci) / (double)((float)n - 0.5f));
    double pl = 0, pl1 = 0, pl2 = 0, dpl, d2pl, xp;
    int it = 0;
    do {
      pl1 = 1.0;
      pl = x;
      for (int l = 2; l <= n1; l++) {
        pl2 = pl1;
        pl1 = pl;
        pl = ((double)(2 * l - 1) * x * pl1 - (double)(l - 1) * pl2) / (double)l;
      }
      dpl = (double)n1 * (x * pl - pl1) / (x * x - 1.0);
      d2pl = (2.0 * x * dpl - (double)(n1 * (n1 + 1)) * pl) / (1.0 - x * x);
      xp = x;
      x = xp - dpl / d2pl;
      it++;
    } while (std::fabs(x - xp) > 3.0e-14 && it < 10);
    mu[j - 1] = x;
    w[j - 1] = 2.0 / ((double)(n * n1) * pl * pl);
  }
  mu[num - 1] = 1.0;
  w[num - 1] = 2.0 / (double)(n * n1);
}

static const char *err_text(int32_t code) {
  switch (code & ~0xff) {
    case E_NLYRS: return "nlyrs out of range";
    case E_NAN_LEVEL: return "found nan in level variables";
    case E_NAN_LAYER: return "found nan in layer variables";
    case E_HGT_ORDER: return "atmo_hgt_lev must be defined bottom-up";
    case E_GAS: return "error in rosen98_gasabs (frequency, temperature or pressure out of range)";
    case E_SURFACE: return "surface model not supported (land atlases)";
    case E_MASS_SIZE: return "error in make_mass_size";
    case E_DIST_PARAMS: return "error in make_dist_params";
    case E_ADAPTIVE: return "make_dist: adaptive grid failed";
    case E_MOMENT: return "calc_moment: invalid distribution";
    case E_SOFT: return "make_soft_spheroid: invalid density (frequency dependent density unsupported)";
    case E_MIE_INPUT: return "calc_mie_spheres: invalid bin input";
    case E_MIE_TERMS: return "calc_mie_spheres: Mie series too long for the device kernel";
    case E_MIE_RESULT: return "calc_mie_spheres: extinction/scatter/back_scatt must be positive";
    case E_KEXT: return "rt_kexttot smaller than zero or nan";
    case E_NLEG: return "nlegen_coef greater maxnleg";
    case E_NORM: return "Scattering function not normalized";
    case E_SINGULAR: return "Matrix inversion failed";
    case E_REFL: return "Wrong or no ground_type selected";
    case E_NAN_COEF: return "nan in Legendre coefficients";
    case E_SSRG_INPUT:
      return "calc_self_similar_rayleigh_gans_rt3: assertion error (liq_ice must be -1, canting 0 or 90, "
             "as_ratio > 0)";
    case E_TMATRIX:
      return "error in calc_tmatrix (as_ratio must be > 0; T-matrix order or Gauss points above the "
             "device limits)";
    default: return "error";
  }
}

// ------------------------------------------------------------------
// small kernels of the scheduler
// ------------------------------------------------------------------

// merge column and (column, channel) status into the result status array
__global__ void k_status(Dims d, Work w, int32_t *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.ncol * d.nf) return;
  int col = i / d.nf;
  int32_t a = w.col_status[col], b = w.cf_status[i];
  int32_t v = ((a & 0xff) >= (b & 0xff)) ? a : b;
  w.cf_status[i] = v;
  if (out) out[i] = v & 0xff;
}

// per-layer diagnostics of one chunk (bottom-up layers)
__global__ void k_diag(Dims d, Work w, double *kextatmo, double *kexttot, double *salbedo,
                       int32_t *nlegen, int32_t *ndoubles) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t n = (size_t)d.ncol_chunk * d.ML * d.nf;
  if (i >= n) return;
  int f = i % d.nf;
  int z = (i / d.nf) % d.ML;
  int col = d.col0 + (int)(i / ((size_t)d.nf * d.ML));
  size_t o = ((size_t)col * d.ML + z) * d.nf + f;
  bool in = z < w.nl[col];
  int item = in ? w.layer_item[(size_t)col * d.ML + z] : -1;
  if (kextatmo) kextatmo[o] = in ? w.kgas[((size_t)col * d.nf + f) * d.ML + z] : __builtin_nan("");
  if (item >= 0) {
    size_t itf = (size_t)(item - d.item0) * d.nf + f;
    if (kexttot) kexttot[o] = w.kext[itf];
    if (salbedo) salbedo[o] = w.salb[itf];
    if (nlegen) nlegen[o] = w.nleg[itf] > 0 ? w.nleg[itf] : -1;
    if (ndoubles) ndoubles[o] = w.ndbl[itf];
  } else {
    if (kexttot) kexttot[o] = in ? 0.0 : __builtin_nan("");
    if (salbedo) salbedo[o] = in ? 0.0 : __builtin_nan("");
    if (nlegen) nlegen[o] = in ? -1 : -9999;
    if (ndoubles) ndoubles[o] = in ? -1 : -9999;
  }
}

// ------------------------------------------------------------------
// C-ABI
// ------------------------------------------------------------------

extern "C" const char *pam_version(void) { return "pamtra_amd 0.1.0 (gfx950)"; }

extern "C" int32_t pam_nummu(void) { return PAM_NUMMU; }

extern "C" int pam_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

extern "C" int pam_context_create(pam_context **out, int device) {
  if (!out) return PAM_FATAL;
  *out = nullptr;
  int n = pam_device_count();
  if (device < 0 || device >= n) return PAM_FATAL;
  if (hipSetDevice(device) != hipSuccess) return PAM_FATAL;
  pam_context *c = new pam_context();
  c->device = device;
  lobatto(NMU, c->mu, c->qw);
  if (hipMalloc(&c->d_mu, 2 * NMU * sizeof(double)) != hipSuccess) {
    delete c;
    return PAM_FATAL;
  }
  c->d_qw = c->d_mu + NMU;
  double tmp[2 * NMU];
  memcpy(tmp, c->mu, sizeof(c->mu));
  memcpy(tmp + NMU, c->qw, sizeof(c->qw));
  if (hipMemcpy(c->d_mu, tmp, sizeof(tmp), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(c->d_mu);
    delete c;
    return PAM_FATAL;
  }
  {
    AngleTables *t = device_tables(device);
    if (!t || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      (void)hipFree(c->d_mu);
      delete c;
      return PAM_FATAL;
    }
    if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess) {
      (void)hipStreamDestroy(c->stream);
      (void)hipFree(c->d_mu);
      delete c;
      return PAM_FATAL;
    }
    c->d_atab = t->tab;
    c->d_atab_f = t->tab_f;
    c->d_atab_off = t->off;
    c->atab_nmax = t->nmax;
    c->d_ptab = t->ptab;
    c->d_ptab_f = t->ptab_f;
    c->d_ptab_off = t->poff;
    c->d_tm_gauss = t->tm_gauss;
  }
  // per-call workspace budget: half of the free HBM, at most 96 GiB
  // (PAM_WORKSPACE_GB, read at each call, overrides it)
  size_t freeb = 0, totalb = 0;
  double gb = 32.0;
  if (hipMemGetInfo(&freeb, &totalb) == hipSuccess && freeb > 0)
    gb = std::min(96.0, 0.5 * (double)freeb / (double)(1ull << 30));
  c->budget_bytes = (size_t)(gb * (1ull << 30));
  if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    c->ncu = 0;
  c->tune[4] = PIPE_DEFAULT;
  c->scratch = scratch_address(device);
  if (!c->scratch) {
    (void)hipStreamDestroy(c->stream2);
    (void)hipStreamDestroy(c->stream);
    (void)hipFree(c->d_mu);
    delete c;
    return PAM_FATAL;
  }
  *out = c;
  return PAM_SUCCESS;
}

extern "C" int pam_calc_tmatrix(pam_context *c, double freq_hz, double mre, double mim,
                                int32_t phase, int32_t nbins, const double *dmax,
                                const double *del_d, const double *ndens, const double *density,
                                const double *as_ratio, const double *canting, double *smat,
                                double *ext, double *emis, int32_t *nmax) {
  if (!c || nbins <= 0 || !dmax || !del_d || !ndens || !density || !as_ratio || !canting || !smat ||
      !ext || !emis || !(freq_hz > 0.0))
    return PAM_FATAL;
  if (hipSetDevice(c->device) != hipSuccess) return PAM_FATAL;
  const size_t nacc = tm_nacc(), nb = (size_t)nbins;
  const size_t bytes = (tm_slot_doubles() + tm_geo_doubles() + nacc + 6 * nb) * 8 + 8 + 4 * nb;
  void *buf = nullptr;
  if (hipMalloc(&buf, bytes) != hipSuccess) return PAM_FATAL;
  double *scr = (double *)buf, *geo = scr + tm_slot_doubles(), *out = geo + tm_geo_doubles();
  double *in = out + nacc;
  int *st = (int *)(in + 6 * nb), *nm = st + 2;
  int rc = PAM_FATAL;
  const double *src[6] = {dmax, del_d, ndens, density, as_ratio, canting};
  bool ok = true;
  for (int k = 0; k < 6 && ok; k++)
    ok = hipMemcpyAsync(in + k * nb, src[k], nb * 8, hipMemcpyHostToDevice, c->stream) ==
         hipSuccess;
  std::vector<int> zeros(nb, 0);
  if (ok) ok = hipMemcpyAsync(nm, zeros.data(), 4 * nb, hipMemcpyHostToDevice, c->stream) == hipSuccess;
  if (ok) {
    launch_tm_hydro(nbins, in, in + nb, in + 2 * nb, in + 3 * nb, in + 4 * nb, in + 5 * nb, freq_hz,
                    mre, mim, phase, c->d_mu, c->d_qw, c->d_tm_gauss, geo, scr, out, st, nm,
                    c->stream);
    std::vector<double> h(nacc);
    int hst = -1;
    std::vector<int> hnm(nb);
    ok = hipGetLastError() == hipSuccess &&
         hipMemcpyAsync(h.data(), out, nacc * 8, hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
         hipMemcpyAsync(&hst, st, 4, hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
         hipMemcpyAsync(hnm.data(), nm, 4 * nb, hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
         hipStreamSynchronize(c->stream) == hipSuccess;
    if (ok && hst == 0) {
      const size_t ns = (size_t)TM_NSMAT;
      memcpy(smat, h.data(), ns * 8);
      memcpy(ext, h.data() + ns, 4 * NMU * 8);
      memcpy(emis, h.data() + ns + 4 * NMU, 2 * NMU * 8);
      if (nmax)
        for (size_t i = 0; i < nb; i++) nmax[i] = hnm[i];
      rc = PAM_SUCCESS;
    }
  }
  (void)hipFree(buf);
  return rc;
}

extern "C" int pam_context_destroy(pam_context *c) {
  if (!c) return PAM_SUCCESS;
  (void)hipSetDevice(c->device);
  c->colbuf.release();
  c->itembuf.release();
  c->iobuf.release();
  c->smallbuf.release();
  c->rtbuf.release();
  c->miebuf.release();
  c->tmbuf.release();
  drop_graph(c);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->d_mu) (void)hipFree(c->d_mu);
  delete c;
  return PAM_SUCCESS;
}

extern "C" const char *pam_stage_names(void) { return k_stage_names; }

extern "C" int pam_last_stage_times(const pam_context *c, double *ms, int32_t n) {
  if (!c || !ms) return PAM_FATAL;
  for (int i = 0; i < n && i < NSTAGE; i++) ms[i] = c->stage_ms[i];
  return PAM_SUCCESS;
}

extern "C" int pam_last_work(const pam_context *c, int64_t *counts, int32_t n) {
  if (!c || !counts) return PAM_FATAL;
  for (int i = 0; i < n && i < PAM_NWORK; i++) counts[i] = c->work[i];
  return PAM_SUCCESS;
}

extern "C" const char *pam_tuning_keys(void) { return k_tuning_keys; }

extern "C" int pam_context_set_tuning(pam_context *c, const char *key, int64_t value) {
  const int i = tuning_index(key);
  if (!c || i < 0 || value < 0 || value > 64) return PAM_FATAL;
  if (i == 3 && value > 1) return PAM_FATAL;
  if (i == 4 && value > PIPE_MAX) return PAM_FATAL;
  c->tune[i] = value;
  if (i == 3 && !value) drop_graph(c);
  return PAM_SUCCESS;
}

extern "C" int64_t pam_context_get_tuning(const pam_context *c, const char *key) {
  const int i = tuning_index(key);
  if (!c || i < 0) return -1;
  return c->tune[i];
}

extern "C" int pam_context_set_workspace_limit(pam_context *c, int64_t bytes) {
  if (!c || bytes <= 0) return PAM_FATAL;
  c->budget_bytes = (size_t)bytes;
  return PAM_SUCCESS;
}

extern "C" int64_t pam_context_bytes(const pam_context *c) {
  if (!c) return -1;
  return (int64_t)(c->colbuf.cap + c->itembuf.cap + c->iobuf.cap + c->smallbuf.cap + c->rtbuf.cap +
                   c->miebuf.cap);
}

extern "C" int pam_context_trim(pam_context *c) {
  if (!c) return PAM_FATAL;
  if (hipSetDevice(c->device) != hipSuccess) return PAM_FATAL;
  if (c->stream && hipStreamSynchronize(c->stream) != hipSuccess) return PAM_FATAL;
  drop_graph(c);
  c->colbuf.release();
  c->itembuf.release();
  c->iobuf.release();
  c->smallbuf.release();
  c->rtbuf.release();
  c->miebuf.release();
  c->tmbuf.release();
  return PAM_SUCCESS;
}

// carve aligned sub-arrays out of one allocation
struct Carver {
  char *base;
  size_t off = 0;
  explicit Carver(void *b) : base((char *)b) {}
  template <class T>
  T *take(size_t n) {
    off = (off + 255) & ~(size_t)255;
    T *p = (T *)(base + off);
    off += n * sizeof(T);
    return p;
  }
};

static size_t col_bytes(const Dims &d) {
  size_t n = 0;
  auto add = [&](size_t b) { n += ((b + 255) & ~(size_t)255); };
  size_t ncol = d.ncol, ML = d.ML, nf = d.nf;
  add(2 * ncol * (ML + 1) * 8);
  add(6 * ncol * ML * 8);
  add(ncol * 8);
  add(ncol * 4);
  add(ncol * d.nout * 4);
  add(ncol * 4);
  add(ncol * ML * 4);
  add(ncol * 4);
  add(ncol * 4);
  add(ncol * ML * 4);
  add(ncol * nf * ML * 8);
  add(2 * ncol * nf * 2 * NMU * 8);
  add(ncol * nf * 4);
  add(ncol * nf * 4);
  add(ncol * nf * 4);
  add(ncol * nf * 4);
  add(ncol * nf * 4);
  add(8 * 8);
  for (int i = 0; i < 16; i++) add(0);
  return n + 256 * 32;
}

static size_t item_bytes_per(const Dims &d) {
  size_t per = 2 * 4 + (size_t)d.nh * d.nbmax * 4 * 8 + (size_t)d.nh * (4 + 8);
  per += (size_t)d.nf * (3 * 8 + 4 + NCOEF * MAXNLEG * 8 + 4 + 4 + 2 * NN2 * 8 + 2 * NN * 8 + 4 + 8 +
                         4 * 8 + 4 * NMU * 8);
  return per;
}

static void carve_items(Work &w, void *base, const Dims &d, size_t nitems) {
  Carver cv(base);
  size_t nif = nitems * d.nf;
  w.item_col = cv.take<int32_t>(nitems);
  w.item_layer = cv.take<int32_t>(nitems);
  size_t nb = nitems * d.nh * d.nbmax;
  w.dsd_d = cv.take<double>(nb);
  w.dsd_dd = cv.take<double>(nb);
  w.dsd_nd = cv.take<double>(nb);
  w.dsd_rho = cv.take<double>(nb);
  w.dsd_flag = cv.take<int32_t>(nitems * d.nh);
  w.dsd_ar = cv.take<double>(nitems * d.nh);
  w.kext = cv.take<double>(nif);
  w.salb = cv.take<double>(nif);
  w.back = cv.take<double>(nif);
  w.nleg = cv.take<int32_t>(nif);
  w.coef = cv.take<double>(nif * NCOEF * MAXNLEG);
  w.ndbl = cv.take<int32_t>(nif);
  w.itf_status = cv.take<int32_t>(nif);
  w.R = cv.take<double>(nif * NN2);
  w.T = cv.take<double>(nif * NN2);
  w.Sp = cv.take<double>(nif * NN);
  w.Sm = cv.take<double>(nif * NN);
  w.redo_dbl = cv.take<int32_t>(nif);
  w.ph_rank = cv.take<int32_t>(nif);
  w.ph_perm = cv.take<int32_t>(nif);
  w.mie_m = cv.take<double>(nif * 4);
  w.tm_ext = cv.take<double>(nif * 2 * NMU);
  w.tm_emis = cv.take<double>(nif * 2 * NMU);
}

extern "C" int64_t pam_workspace_bytes(int64_t ncol, int32_t max_nlyrs, int32_t nhydro,
                                       int32_t nbin_max, int32_t nfreq, int32_t noutlevels,
                                       int64_t nitems) {
  if (ncol <= 0 || max_nlyrs <= 0 || nhydro < 0 || nfreq <= 0 || noutlevels <= 0 || nitems < 0)
    return -1;
  Dims d;
  memset(&d, 0, sizeof(d));
  d.ncol = (int)ncol;
  d.ML = max_nlyrs;
  d.nh = nhydro;
  d.nbmax = std::max(2, nbin_max);
  d.nf = nfreq;
  d.nout = noutlevels;
  const size_t c = (size_t)ncol, L = max_nlyrs, nh = nhydro, nf = nfreq, no = noutlevels;
  // pam_run staging: every profile and 4D descriptor array, and every result
  size_t io = c * 4 * 4 + 4 * c * (L + 1) * 8 + 4 * c * L * 8 + 3 * c * L * nh * 8 + 5 * c * 8 +
              c + c * no * 8 + 13 * c * L * nh * 8;
  io += c * no * PAM_NANGLES * nf * 2 * 8 + c * 2 * nf * NMU * 8 + PAM_NANGLES * 8 + c * nf * 4 +
        c * L * nf * (3 * 8 + 2 * 4);
  const size_t small = 4096 + sizeof(pam_hydro_desc) * std::max<size_t>(1, nh) + 16 * nf + 256 + 64 * 8 +
                       3 * 4 * (PH_NBUCKET + 1) + 256;
  return (int64_t)(col_bytes(d) + (size_t)std::max<int64_t>(1, nitems) * item_bytes_per(d) +
                   64 * 256 + io + 64 * 256 + small);
}

// One call.  arena null: the workspace grows in the context's buffers, stage
// times are recorded with events, and a single-chunk call may replay graphs.
// arena set: every workspace region is carved from it, the items are chunked
// to fit what is left, and nothing is allocated or created (no events, no
// graphs); a call that cannot fit is refused.
static int run_device(pam_context *c, const pam_settings *st, const pam_hydro_desc *hydros,
                      const pam_hydro_4d *h4in, const pam_profile *pin, const double *freqs,
                      int32_t nfreq, pam_results *res, hipStream_t stream,
                      Arena *arena = nullptr) {
  if (!c || !st || !pin || !freqs || !res || !res->tb) {
    set_msg(res, "pam_run: NULL argument");
    return PAM_FATAL;
  }
  if (!st->passive) {
    set_msg(res, "passive=False is not supported");
    return PAM_FATAL;
  }
  if (st->active && (st->radar_npol < 1 || st->radar_npol > 5 || st->radar_attenuation < 0 ||
                     st->radar_attenuation > 2)) {
    set_msg(res, "invalid radar settings");
    return PAM_FATAL;
  }
  if (st->gas_mod != 0 || st->liq_mod != 0) {
    set_msg(res, "only gas_mod R98 and liq_mod TKC are implemented");
    return PAM_FATAL;
  }
  if (pin->nhydro > MAX_HYDRO || nfreq <= 0 || pin->ngridx <= 0 || pin->ngridy <= 0 ||
      pin->max_nlyrs <= 0 || pin->noutlevels <= 0) {
    set_msg(res, "pam_run: invalid dimensions");
    return PAM_FATAL;
  }
  for (int i = 1; i < nfreq; i++)
    if (!(freqs[i] >= freqs[i - 1])) {
      set_msg(res, "frequencies must be sorted ascending");
      return PAM_FATAL;
    }
  HIP_OK(hipSetDevice(c->device));
  // a workspace region: from the arena, or the context buffer grown to n
  auto region = [&](DevBuf &b, size_t n) -> void * {
    if (arena) return arena->take(n);
    return b.ensure(n) == hipSuccess ? b.p : nullptr;
  };
  auto no_room = [&](const char *what) -> int {
    char m[200];
    if (arena)
      snprintf(m, sizeof(m),
               "pam_run_device_noalloc: %s does not fit the %zu MiB scratch; split the call",
               what, (size_t)(arena->size >> 20));
    else
      snprintf(m, sizeof(m), "HIP error: device allocation failed (%s)", what);
    set_msg(res, m);
    return PAM_FATAL;
  };
  if (!stream && c->tune[3] && !arena) {
    // the legacy stream cannot be captured: the call runs on the context's
    // stream, after everything queued before it on the legacy stream (the
    // call synchronises before it returns)
    hipEvent_t e;
    HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_OK(hipEventRecord(e, nullptr));
    HIP_OK(hipStreamWaitEvent(c->stream, e, 0));
    (void)hipEventDestroy(e);
    stream = c->stream;
  }
  for (int i = 0; i < NSTAGE; i++) c->stage_ms[i] = 0.0;
  for (int i = 0; i < PAM_NWORK; i++) c->work[i] = 0;

  Dims d;
  memset(&d, 0, sizeof(d));
  d.nx = pin->ngridx;
  d.ny = pin->ngridy;
  d.ncol = d.nx * d.ny;
  d.ML = pin->max_nlyrs;
  d.nout = pin->noutlevels;
  d.nh = pin->nhydro;
  d.nf = nfreq;
  d.nbmax = 1;
  for (int h = 0; h < d.nh; h++) {
    const int dn = hydros[h].dist_name;
    int nb = (dn == PAM_DIST_MONO || dn == PAM_DIST_MONO_COSMO_ICE) ? 2 : hydros[h].nbin;
    d.nbmax = std::max(d.nbmax, nb);
  }
  if (h4in && h4in->nbin) {
    std::vector<int32_t> hb((size_t)d.ncol * d.ML * d.nh);
    HIP_OK(hipMemcpy(hb.data(), h4in->nbin, hb.size() * 4, hipMemcpyDeviceToHost));
    for (int32_t v : hb) d.nbmax = std::max(d.nbmax, (int)v);
  }
  if (h4in && h4in->fs_nbin > 0) {
    const double *const fs[] = {h4in->fs_rho,     h4in->fs_d,         h4in->fs_d_bound,
                                h4in->fs_n,       h4in->fs_mass,      h4in->fs_as_ratio,
                                h4in->fs_canting, h4in->fs_rg_kappa,  h4in->fs_rg_beta,
                                h4in->fs_rg_gamma, h4in->fs_rg_zeta};
    for (const double *a : fs)
      if (!a) {
        set_msg(res, "hydro_fullspec: every full-spectrum array is required");
        return PAM_FATAL;
      }
    d.nbmax = std::max(d.nbmax, (int)h4in->fs_nbin);
  }
  if (d.nbmax > PAM_MAX_NBIN) {
    set_msg(res, "nbin: at most 4000 size bins per hydrometeor are supported");
    return PAM_FATAL;
  }
  {
    // the T-matrix producer: not beside the Legendre-series models (the
    // reference's layer albedo then reads an unset value), not with the radar
    bool tm = false, series = false;
    for (int i = 0; i < d.nh; i++) {
      if (hydros[i].scat_name == PAM_SCAT_TMATRIX) {
        tm = true;
      } else if (hydros[i].scat_name != PAM_SCAT_DISABLED) {
        series = true;
      }
    }
    if (tm && series) {
      set_msg(res, "tmatrix cannot be mixed with mie-sphere or ssrg-rt3 hydrometeors");
      return PAM_FATAL;
    }
    if (tm && st->active) {
      set_msg(res, "tmatrix: radar simulation is not implemented");
      return PAM_FATAL;
    }
  }
  d.col0 = 0;
  d.ncol_chunk = d.ncol;
  d.nitems = 0;
  d.item0 = 0;

  Settings s;
  memset(&s, 0, sizeof(s));
  s.lgas = st->lgas_extinction;
  s.lhyd_abs = st->lhyd_absorption;
  s.lhyd_scat = st->lhyd_scattering;
  s.lhyd_emis = st->lhyd_emission;
  s.incl_hydro_rho = st->hydro_includehydroinrhoair;
  s.limit_density = st->hydro_limit_density_area;
  s.adaptive = st->hydro_adaptive_grid;
  s.conserve = st->conserve_mass_rescale_dsd;
  s.outpol = st->outpol;
  s.precision = st->precision;
  s.threshold = st->hydro_threshold;
  s.softsphere_min = st->hydro_softsphere_min_density;
  s.emissivity = st->emissivity;
  {
    const char *inv = getenv("PAM_INVERSE");
    s.inverse_gj = !inv ? 0 : strcmp(inv, "gauss-jordan") == 0 ? 1 : strcmp(inv, "queue-all") == 0 ? 2 : 0;
  }
  s.active = st->active ? 1 : 0;
  s.radar_att = st->radar_attenuation;
  s.radar_npol = st->active ? st->radar_npol : 0;
  for (int i = 0; i < 5; i++) s.radar_pol[i] = st->radar_pol[i];
  s.pers_mie = (int)c->tune[0];
  s.pers_dbl = (int)c->tune[1];
  s.pers_add = (int)c->tune[2];
  s.ncu = c->ncu;
  for (int i = 0; i < d.nh; i++) {
    if (hydros[i].scat_name == PAM_SCAT_SSRG_RT3) s.ssrg = 1;
    // 2: some T-matrix hydrometeor has a canting angle
    if (hydros[i].scat_name == PAM_SCAT_TMATRIX)
      s.tmatrix = std::max(s.tmatrix, hydros[i].canting > 0.0 ? 2 : 1);
  }
  {
    const char *mk = getenv("PAM_MIE");
    s.mie_chunked = (mk && strcmp(mk, "chunked") == 0) ? 1 : 0;
    s.mie_fp64 = (mk && strcmp(mk, "fp64") == 0) ? 1 : 0;
    s.fault = FAULT_NONE;
    if (const char *ft = getenv("PAM_TEST_FAULT")) {
      char kind[16] = {0};
      int fc = -1, ff = -1;
      if (sscanf(ft, "%15[a-z_]:%d:%d", kind, &fc, &ff) == 3) {
        s.fault = strcmp(kind, "nan_coef") == 0 ? FAULT_NAN_COEF
                  : strcmp(kind, "singular") == 0 ? FAULT_SINGULAR : FAULT_NONE;
        s.fault_col = fc;
        s.fault_ch = ff;
      }
    }
  }

  Profile p;
  p.nlyrs = pin->nlyrs;
  p.hgt_lev = pin->hgt_lev;
  p.temp_lev = pin->temp_lev;
  p.press_lev = pin->press_lev;
  p.relhum_lev = pin->relhum_lev;
  p.hgt = pin->hgt;
  p.temp = pin->temp;
  p.press = pin->press;
  p.relhum = pin->relhum;
  p.hydro_q = pin->hydro_q;
  p.hydro_n = pin->hydro_n;
  p.hydro_reff = pin->hydro_reff;
  p.groundtemp = pin->groundtemp;
  p.wind10u = pin->wind10u;
  p.wind10v = pin->wind10v;
  p.sfc_salinity = pin->sfc_salinity;
  p.sfc_emissivity = pin->sfc_emissivity;
  p.sfc_type = pin->sfc_type;
  p.sfc_model = pin->sfc_model;
  p.sfc_refl = pin->sfc_refl;
  p.obs_height = pin->obs_height;
  Hydro4D h4;
  memset(&h4, 0, sizeof(h4));
  if (h4in) {
    h4.as_ratio = h4in->as_ratio;
    h4.rho_ms = h4in->rho_ms;
    h4.a_ms = h4in->a_ms;
    h4.b_ms = h4in->b_ms;
    h4.alpha_as = h4in->alpha_as;
    h4.beta_as = h4in->beta_as;
    h4.p_1 = h4in->p_1;
    h4.p_2 = h4in->p_2;
    h4.p_3 = h4in->p_3;
    h4.p_4 = h4in->p_4;
    h4.d_1 = h4in->d_1;
    h4.d_2 = h4in->d_2;
    h4.nbin = h4in->nbin;
    h4.fs_nbin = h4in->fs_nbin > 0 ? h4in->fs_nbin : 0;
    h4.fs_rho = h4in->fs_rho;
    h4.fs_d = h4in->fs_d;
    h4.fs_d_bound = h4in->fs_d_bound;
    h4.fs_n = h4in->fs_n;
    h4.fs_mass = h4in->fs_mass;
    h4.fs_as_ratio = h4in->fs_as_ratio;
    h4.fs_canting = h4in->fs_canting;
    h4.fs_rg_kappa = h4in->fs_rg_kappa;
    h4.fs_rg_beta = h4in->fs_rg_beta;
    h4.fs_rg_gamma = h4in->fs_rg_gamma;
    h4.fs_rg_zeta = h4in->fs_rg_zeta;
  }

  // small buffers: hydro table, frequencies, counters
  size_t small = 4096 + sizeof(pam_hydro_desc) * std::max(1, d.nh) + 16 * (size_t)nfreq + 256 + 64 * 8 +
                 2 * 3 * 4 * (PH_NBUCKET + 1) + 1024;
  void *small_p = region(c->smallbuf, small);
  if (!small_p) return no_room("the call's tables");
  Carver sc(small_p);
  pam_hydro_desc *d_hyd = sc.take<pam_hydro_desc>(std::max(1, d.nh));
  double *d_freqs = sc.take<double>(nfreq);
  unsigned long long *d_cnt = sc.take<unsigned long long>(PAM_NWORK);
  // per stream of a pipelined call: redo counts, work queues, phase buckets
  uint32_t *d_redo_n = sc.take<uint32_t>(2 * 3);
  uint32_t *d_queue = sc.take<uint32_t>(2 * 4);
  double *d_k2 = sc.take<double>(nfreq);
  uint32_t *d_ph = sc.take<uint32_t>(2 * 3 * (PH_NBUCKET + 1));
  if (d.nh > 0)
    HIP_OK(hipMemcpyAsync(d_hyd, hydros, sizeof(pam_hydro_desc) * d.nh, hipMemcpyHostToDevice,
                          stream));
  HIP_OK(hipMemcpyAsync(d_freqs, freqs, 8 * (size_t)nfreq, hipMemcpyHostToDevice, stream));
  if (st->active && st->radar_k2)
    HIP_OK(hipMemcpyAsync(d_k2, st->radar_k2, 8 * (size_t)nfreq, hipMemcpyHostToDevice, stream));
  // counters and status words are cleared by a fill kernel, not a memset
  launch_fill_u32((uint32_t *)d_cnt, 2 * PAM_NWORK, 0, stream);

  // per-column workspace
  void *col_p = region(c->colbuf, col_bytes(d));
  if (!col_p) return no_room("the per-column workspace");
  Work w;
  memset(&w, 0, sizeof(w));
  int32_t *redo_add2 = nullptr, *redo_lev2 = nullptr;
  {
    Carver cv(col_p);
    size_t ncol = d.ncol, ML = d.ML, nf = d.nf;
    w.hgt_lev = cv.take<double>(ncol * (ML + 1));
    w.temp_lev = cv.take<double>(ncol * (ML + 1));
    w.temp = cv.take<double>(ncol * ML);
    w.press = cv.take<double>(ncol * ML);
    w.rho_v = cv.take<double>(ncol * ML);
    w.q_hum = cv.take<double>(ncol * ML);
    w.dz = cv.take<double>(ncol * ML);
    w.hgt = cv.take<double>(ncol * ML);
    w.groundtemp = cv.take<double>(ncol);
    w.nl = cv.take<int32_t>(ncol);
    w.outlev = cv.take<int32_t>(ncol * d.nout);
    w.col_status = cv.take<int32_t>(ncol);
    w.hmask = cv.take<uint32_t>(ncol * ML);
    w.col_nitems = cv.take<int32_t>(ncol);
    w.col_item0 = cv.take<int32_t>(ncol);
    w.layer_item = cv.take<int32_t>(ncol * ML);
    w.kgas = cv.take<double>(ncol * nf * ML);
    w.sfc_emis = cv.take<double>(ncol * nf * 2 * NMU);
    w.sfc_refl = cv.take<double>(ncol * nf * 2 * NMU);
    w.cf_status = cv.take<int32_t>(ncol * nf);
    w.redo_add = cv.take<int32_t>(ncol * nf);
    w.redo_lev = cv.take<int32_t>(ncol * nf);
    redo_add2 = cv.take<int32_t>(ncol * nf);
    redo_lev2 = cv.take<int32_t>(ncol * nf);
    launch_fill_u32((uint32_t *)w.col_status, ncol, 0, stream);
    launch_fill_u32((uint32_t *)w.cf_status, ncol * nf, 0, stream);
  }
  w.mu = c->d_mu;
  w.qw = c->d_qw;
  w.atab = c->d_atab;
  w.atab_f = c->d_atab_f;
  w.atab_off = c->d_atab_off;
  w.atab_nmax = c->atab_nmax;
  w.ptab = c->d_ptab;
  w.ptab_f = c->d_ptab_f;
  w.tm_gauss = c->d_tm_gauss;
  w.ptab_off = c->d_ptab_off;
  w.ph_cnt = d_ph;
  w.ph_off = d_ph + (PH_NBUCKET + 1);
  w.ph_toff = d_ph + 2 * (PH_NBUCKET + 1);
  w.counters = d_cnt;
  w.redo_n = d_redo_n;
  w.queue = d_queue;
  w.radar_k2 = (st->active && st->radar_k2) ? d_k2 : nullptr;

  std::vector<hipEvent_t> evs;
  std::vector<int> ev_stage;
  auto mark = [&](int stage) -> hipError_t {
    if (arena) return hipSuccess;  // no stage times without events
    hipEvent_t e;
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) return r;
    evs.push_back(e);
    ev_stage.push_back(stage);
    return hipEventRecord(e, stream);
  };
  auto free_events = [&]() {
    for (auto e : evs) (void)hipEventDestroy(e);
    evs.clear();
  };

  HIP_OK(mark(-1));
  launch_columns(d, s, p, d_hyd, h4, d_freqs, w, stream);
  launch_surface(d, s, p, d_freqs, w, res->emissivity, stream);
  HIP_OK(hipGetLastError());
  HIP_OK(mark(0));

  // scattering layers per column -> item offsets (host scan) and chunks
  std::vector<int32_t> cnt(d.ncol), off(d.ncol);
  HIP_OK(hipMemcpyAsync(cnt.data(), w.col_nitems, 4 * (size_t)d.ncol, hipMemcpyDeviceToHost,
                        stream));
  HIP_OK(hipStreamSynchronize(stream));
  int64_t total = 0;
  for (int i = 0; i < d.ncol; i++) {
    off[i] = (int32_t)total;
    total += cnt[i];
  }
  if (total > 0x7fffffff) {
    free_events();
    set_msg(res, "too many scattering layers in one call");
    return PAM_FATAL;
  }
  HIP_OK(hipMemcpyAsync(w.col_item0, off.data(), 4 * (size_t)d.ncol, hipMemcpyHostToDevice,
                        stream));
  // Pipelining (tuning key "pipeline" = k >= 2, not in the allocation-free
  // entry): at least k column chunks on two streams with a workspace each;
  // chunk c+1 starts once chunk c's Mie stage is done, so that its DSD and
  // Mie run beside chunk c's phase matrix, doubling and adding.
  const int64_t total_units = total * d.nf;
  int64_t npipe = 1;
  if (!arena && c->tune[4] >= 2)
    npipe = std::min<int64_t>(c->tune[4], total_units / PIPE_MIN_UNITS);
  const bool pipelined = npipe >= 2;
  const int nlane = pipelined ? 2 : 1;
  const size_t per_item = item_bytes_per(d);
  size_t budget = c->budget_bytes;
  if (const char *e = getenv("PAM_WORKSPACE_GB")) budget = (size_t)(atof(e) * (1ull << 30));
  size_t max_items = std::max<size_t>(1, budget / per_item / nlane);
  // the Mie kernel's U slots: one per workgroup of its persistent grid;
  // within the fixed scratch of the allocation-free entry, at most a quarter
  // of it
  const bool mie_slotted = !s.ssrg && !s.mie_chunked;
  size_t mie_nslots = 0;
  if (mie_slotted) {
    const int per_cu = s.pers_mie > 0 ? s.pers_mie : MIE2_PER_CU;
    w.mie_ld = mie_slot_rows(d.nbmax);
    w.mie_slot_doubles = mie_slot_doubles(d.nbmax);
    mie_nslots = (size_t)std::max(1, c->ncu) * per_cu;
    if (arena) mie_nslots = std::min(mie_nslots, arena->size / 4 / (w.mie_slot_doubles * 8));
    if (mie_nslots == 0) return no_room("one Mie workgroup's slot");
  }
  const size_t rt_bytes = nlane * RT_SLOT_DOUBLES * 8 * (size_t)std::max(1, c->ncu);
  const size_t mie_bytes = nlane * mie_nslots * (mie_slotted ? w.mie_slot_doubles * 8 : 0);
  // the T-matrix producer's workgroup slots and geometry table
  size_t tm_nslots = 0;
  if (s.tmatrix) {
    tm_nslots = (size_t)std::max(1, c->ncu) * TM_PER_CU;
    if (arena) tm_nslots = std::min(tm_nslots, arena->size / 4 / (tm_slot_doubles() * 8));
    if (tm_nslots == 0) return no_room("one T-matrix workgroup's slot");
  }
  const size_t tm_lane = (tm_nslots * tm_slot_doubles() + tm_geo_doubles() + 63) / 64 * 64;
  const size_t tm_bytes = s.tmatrix ? nlane * tm_lane * 8 : 0;
  if (arena) {
    // what the optics may take: the arena less the adding scratch and the
    // Mie slots carved after them (each region rounded up to 256 bytes)
    size_t room = arena->left();
    const size_t fixed = (rt_bytes + 255) / 256 * 256 + (mie_bytes + 255) / 256 * 256 +
                         (tm_bytes + 255) / 256 * 256 + 768;
    room = room > fixed ? room - fixed : 0;
    max_items = room > 64 * 256 ? (room - 64 * 256) / per_item : 0;
    if (max_items == 0 && total > 0) return no_room("one scattering layer's optics");
    max_items = std::max<size_t>(1, max_items);
  }
  // chunks of whole columns with about equal item counts (no small tail chunk)
  std::vector<std::pair<int, int>> chunks;
  {
    const int64_t nchunk = std::max<int64_t>(
        npipe, std::max<int64_t>(1, (total + (int64_t)max_items - 1) / (int64_t)max_items));
    const int64_t target = (total + nchunk - 1) / nchunk;
    int c0 = 0;
    int64_t acc = 0;
    for (int i = 0; i < d.ncol; i++) {
      if (acc + cnt[i] > target && i > c0) {
        chunks.push_back({c0, i});
        c0 = i;
        acc = 0;
      }
      acc += cnt[i];
    }
    chunks.push_back({c0, d.ncol});
  }
  size_t chunk_items_max = 0;
  for (auto &ch : chunks) {
    int64_t n = 0;
    for (int i = ch.first; i < ch.second; i++) n += cnt[i];
    chunk_items_max = std::max<size_t>(chunk_items_max, (size_t)n);
  }
  if (arena && chunk_items_max > max_items) return no_room("the optics of one column");
  const size_t item_region =
      (std::max<size_t>(1, chunk_items_max) * per_item + 64 * 256 + 255) & ~(size_t)255;
  void *item_p = region(c->itembuf, nlane * item_region);
  if (!item_p) return no_room("the per-layer optics");
  if (RT_SLOT_DOUBLES > 0) {
    w.rt_scratch = (double *)region(c->rtbuf, rt_bytes);
    if (!w.rt_scratch) return no_room("the adding kernels' scratch");
  }
  if (mie_slotted) {
    w.mie_slots = (int)mie_nslots;
    w.mie_u = (double *)region(c->miebuf, mie_bytes);
    if (!w.mie_u) return no_room("the Mie kernel's slots");
  }
  if (s.tmatrix) {
    w.tm_slots = (int)tm_nslots;
    w.tm_scr = (double *)region(c->tmbuf, tm_bytes);
    if (!w.tm_scr) return no_room("the T-matrix producer's slots");
    w.tm_geo = w.tm_scr + tm_nslots * tm_slot_doubles();
  }
  c->work[0] = d.ncol;
  c->work[1] = d.nf;
  c->work[2] = total * d.nf;
  // the second stream's copy of the workspace description: its own items,
  // queues, redo lists, phase buckets, Mie slots and adding scratch
  Work w2 = w;
  hipStream_t stream2 = pipelined ? c->stream2 : stream;
  if (pipelined) {
    w2.queue = d_queue + 4;
    w2.redo_n = d_redo_n + 3;
    w2.ph_cnt = d_ph + 3 * (PH_NBUCKET + 1);
    w2.ph_off = w2.ph_cnt + (PH_NBUCKET + 1);
    w2.ph_toff = w2.ph_cnt + 2 * (PH_NBUCKET + 1);
    w2.redo_add = redo_add2;
    w2.redo_lev = redo_lev2;
    if (w.mie_u) w2.mie_u = w.mie_u + (size_t)w.mie_slots * w.mie_slot_doubles;
    if (w.tm_scr) {
      w2.tm_scr = w.tm_scr + tm_lane;
      w2.tm_geo = w2.tm_scr + tm_nslots * tm_slot_doubles();
    }
    if (RT_SLOT_DOUBLES > 0)
      w2.rt_scratch = w.rt_scratch + RT_SLOT_DOUBLES * (size_t)std::max(1, c->ncu);
    // stream2 starts after the columns and the surface on the call's stream
    hipEvent_t e;
    HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_OK(hipEventRecord(e, stream));
    HIP_OK(hipStreamWaitEvent(stream2, e, 0));
    (void)hipEventDestroy(e);
  }

  // The kernels of one chunk in five segments, one per timed stage (dsd,
  // mie, phase, doubling, adding).  Counters and status words are cleared by
  // a fill kernel (not a memset), so that each segment can be captured in a
  // graph.
  auto enqueue_segment = [&](int seg, const Dims &dc, Work &wc, hipStream_t sx) -> hipError_t {
    switch (seg) {
      case 1:
        launch_fill_u32(wc.queue, 4, 0, sx);
        launch_fill_u32(wc.redo_n, 3, 0, sx);
        if (dc.nitems > 0)
          launch_fill_u32((uint32_t *)wc.itf_status, (size_t)dc.nitems * d.nf, 0, sx);
        launch_items(dc, wc, sx);
        launch_dsd(dc, s, p, d_hyd, h4, wc, sx);
        break;
      case 2:
        if (s.tmatrix)
          launch_tmatrix(dc, s, d_hyd, h4, d_freqs, wc, sx);
        else
          launch_mie(dc, s, d_hyd, h4, p, d_freqs, wc, sx);
        break;
      case 3:
        // the T-matrix producer writes the phase matrix itself
        if (!s.tmatrix) launch_phase(dc, s, wc, sx);
        break;
      case 4:
        launch_doubling(dc, s, d_freqs, wc, sx);
        break;
      default:
        launch_adding(dc, s, d_freqs, pin->sfc_refl, wc, res->tb, sx);
        if (s.active)
          launch_radar(dc, s, d_freqs, wc, res->ze, res->att_hydro, res->att_atmo, res->radar_hgt,
                       sx);
        if (res->kextatmo || res->kexttot || res->salbedo || res->nlegen || res->ndoubles) {
          size_t nd = (size_t)dc.ncol_chunk * d.ML * d.nf;
          hipLaunchKernelGGL(k_diag, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, sx, dc,
                             wc, res->kextatmo, res->kexttot, res->salbedo, res->nlegen,
                             res->ndoubles);
        }
        break;
    }
    return hipGetLastError();
  };
  const bool use_graph = c->tune[3] && chunks.size() == 1 && !arena;
  // stage times of a pipelined call: each segment from the previous event
  // on its own stream (the segments of the two streams overlap in time)
  int last_ev[2] = {(int)evs.size() - 1, -1};
  std::vector<int> ev_from(evs.size());
  for (size_t i = 0; i < ev_from.size(); i++) ev_from[i] = (int)i - 1;
  auto mark_on = [&](int stage, int lane, hipStream_t sx) -> hipError_t {
    if (arena) return hipSuccess;
    hipEvent_t e;
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) return r;
    evs.push_back(e);
    ev_stage.push_back(stage);
    ev_from.push_back(last_ev[lane]);
    last_ev[lane] = (int)evs.size() - 1;
    return hipEventRecord(e, sx);
  };
  hipEvent_t mie_done = nullptr;  // the previous chunk's Mie stage (pipelined)
  for (size_t ci = 0; ci < chunks.size(); ci++) {
    auto &ch = chunks[ci];
    const int lane = pipelined ? (int)(ci & 1) : 0;
    Work &wc = lane ? w2 : w;
    hipStream_t sx = lane ? stream2 : stream;
    Dims dc;
    memcpy(&dc, &d, sizeof(d));
    dc.col0 = ch.first;
    dc.ncol_chunk = ch.second - ch.first;
    dc.item0 = off[ch.first];
    int64_t n = 0;
    for (int i = ch.first; i < ch.second; i++) n += cnt[i];
    dc.nitems = (int)n;
    carve_items(wc, (char *)item_p + (size_t)lane * item_region, dc, std::max<size_t>(1, (size_t)n));
    if (pipelined) {
      if (mie_done) {
        HIP_OK(hipStreamWaitEvent(sx, mie_done, 0));
        (void)hipEventDestroy(mie_done);
        mie_done = nullptr;
      }
      if (last_ev[lane] < 0 || ci >= 2) HIP_OK(mark_on(-1, lane, sx));
    }
    if (use_graph) {
      // the launch arguments the graphs bake in: shapes, settings,
      // workspace and the caller's device pointers
      std::vector<unsigned char> key(sizeof(Dims) + sizeof(Settings) + sizeof(Work) +
                                     sizeof(Profile) + sizeof(Hydro4D) + 13 * sizeof(void *));
      unsigned char *kp = key.data();
      const void *ptrs[9] = {d_hyd, d_freqs, pin->sfc_refl, res->tb, res->kextatmo,
                             res->kexttot, res->salbedo, res->nlegen, res->ndoubles};
      const void *rptrs[4] = {res->ze, res->att_hydro, res->att_atmo, res->radar_hgt};
      memcpy(kp, &dc, sizeof(Dims));
      kp += sizeof(Dims);
      memcpy(kp, &s, sizeof(Settings));
      kp += sizeof(Settings);
      memcpy(kp, &w, sizeof(Work));
      kp += sizeof(Work);
      memcpy(kp, &p, sizeof(Profile));
      kp += sizeof(Profile);
      memcpy(kp, &h4, sizeof(Hydro4D));
      kp += sizeof(Hydro4D);
      memcpy(kp, ptrs, sizeof(ptrs));
      kp += sizeof(ptrs);
      memcpy(kp, rptrs, sizeof(rptrs));
      if (!(c->gexec[0] && c->gkey == key)) {
        drop_graph(c);
        bool ok = true;
        for (int seg = 1; seg <= 5 && ok; seg++) {
          hipGraph_t g = nullptr;
          ok = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
          if (!ok) break;
          const hipError_t e1 = enqueue_segment(seg, dc, w, stream);
          const hipError_t e2 = hipStreamEndCapture(stream, &g);
          ok = e1 == hipSuccess && e2 == hipSuccess && g;
          if (ok) ok = hipGraphInstantiate(&c->gexec[seg - 1], g, nullptr, nullptr, 0) == hipSuccess;
          if (g) (void)hipGraphDestroy(g);
        }
        (void)hipGetLastError();
        if (ok)
          c->gkey = key;
        else
          drop_graph(c);  // not capturable here: direct launches
      }
    }
    for (int seg = 1; seg <= 5; seg++) {
      if (use_graph && c->gexec[0])
        HIP_OK(hipGraphLaunch(c->gexec[seg - 1], stream));
      else
        HIP_OK(enqueue_segment(seg, dc, wc, sx));
      HIP_OK(mark_on(seg, lane, sx));
      if (pipelined && seg == 2 && ci + 1 < chunks.size()) {
        HIP_OK(hipEventCreateWithFlags(&mie_done, hipEventDisableTiming));
        HIP_OK(hipEventRecord(mie_done, sx));
      }
    }
  }
  if (pipelined) {  // the call's stream waits for the second stream
    hipEvent_t e;
    HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_OK(hipEventRecord(e, stream2));
    HIP_OK(hipStreamWaitEvent(stream, e, 0));
    (void)hipEventDestroy(e);
  }
  hipLaunchKernelGGL(k_status, dim3((unsigned)((d.ncol * d.nf + 255) / 256)), dim3(256), 0,
                     stream, d, w, res->status);
  HIP_OK(hipGetLastError());
  std::vector<int32_t> cf((size_t)d.ncol * d.nf);
  unsigned long long cntv[PAM_NWORK];
  HIP_OK(hipMemcpyAsync(cf.data(), w.cf_status, 4 * cf.size(), hipMemcpyDeviceToHost, stream));
  HIP_OK(hipMemcpyAsync(cntv, d_cnt, sizeof(cntv), hipMemcpyDeviceToHost, stream));
  HIP_OK(hipStreamSynchronize(stream));
  // stage times
  for (size_t i = 1; i < evs.size(); i++) {
    float ms = 0.f;
    const int from = i < ev_from.size() ? ev_from[i] : (int)i - 1;
    if (from >= 0 && ev_stage[i] >= 0 &&
        hipEventElapsedTime(&ms, evs[from], evs[i]) == hipSuccess)
      c->stage_ms[ev_stage[i]] += ms;
  }
  free_events();
  for (int i = 3; i < PAM_NWORK; i++) c->work[i] = (int64_t)cntv[i];
  int worst = PAM_SUCCESS;
  int32_t first = 0;
  for (int32_t v : cf) {
    int sev = v & 0xff;
    if (sev == PAM_INFO) continue;
    if (sev > worst) {
      worst = sev;
      first = v;
    }
  }
  if (worst != PAM_SUCCESS) set_msg(res, err_text(first));
  if (res->angles_deg) {
    double a[PAM_NANGLES];
    for (int j = 0; j < NMU; j++) {
      a[j] = 180.0 - (180. * std::acos(c->mu[NMU - 1 - j]) / PI);
      a[NMU + j] = (180. * std::acos(c->mu[j]) / PI);
    }
    HIP_OK(hipMemcpyAsync(res->angles_deg, a, sizeof(a), hipMemcpyHostToDevice, stream));
    HIP_OK(hipStreamSynchronize(stream));
  }
  return worst;
}

// True when the device can dereference ptr: memory from hipMalloc (torch's
// allocator included), managed memory, or pinned host memory mapped into the
// device's address space.  A plain host pointer reads as unregistered; the
// error that query leaves is cleared.
static bool device_accessible(const void *ptr) {
  hipPointerAttribute_t a;
  memset(&a, 0, sizeof(a));
  const hipError_t e = hipPointerGetAttributes(&a, ptr);
  (void)hipGetLastError();
  if (e != hipSuccess) return false;
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged ||
         (a.type == hipMemoryTypeHost && a.devicePointer != nullptr);
}

// The arrays of a pam_run_device call that the kernels read or write: every
// one must be device memory, or the call is refused before any launch (a
// host array would fault the GPU).  Returns the offending field or NULL.
static const char *first_host_array(const pam_hydro_4d *h4, const pam_profile *p,
                                    const pam_results *r) {
  const struct {
    const void *ptr;
    const char *name;
  } arrays[] = {
      {p->nlyrs, "profile.nlyrs"},           {p->hgt_lev, "profile.hgt_lev"},
      {p->temp_lev, "profile.temp_lev"},     {p->press_lev, "profile.press_lev"},
      {p->relhum_lev, "profile.relhum_lev"}, {p->hgt, "profile.hgt"},
      {p->temp, "profile.temp"},             {p->press, "profile.press"},
      {p->relhum, "profile.relhum"},         {p->hydro_q, "profile.hydro_q"},
      {p->hydro_n, "profile.hydro_n"},       {p->hydro_reff, "profile.hydro_reff"},
      {p->groundtemp, "profile.groundtemp"}, {p->wind10u, "profile.wind10u"},
      {p->wind10v, "profile.wind10v"},       {p->sfc_salinity, "profile.sfc_salinity"},
      {p->sfc_emissivity, "profile.sfc_emissivity"},
      {p->sfc_type, "profile.sfc_type"},     {p->sfc_model, "profile.sfc_model"},
      {p->sfc_refl, "profile.sfc_refl"},     {p->obs_height, "profile.obs_height"},
      {h4 ? h4->as_ratio : nullptr, "hydro4d.as_ratio"},
      {h4 ? h4->rho_ms : nullptr, "hydro4d.rho_ms"},
      {h4 ? h4->a_ms : nullptr, "hydro4d.a_ms"},
      {h4 ? h4->b_ms : nullptr, "hydro4d.b_ms"},
      {h4 ? h4->alpha_as : nullptr, "hydro4d.alpha_as"},
      {h4 ? h4->beta_as : nullptr, "hydro4d.beta_as"},
      {h4 ? h4->p_1 : nullptr, "hydro4d.p_1"},     {h4 ? h4->p_2 : nullptr, "hydro4d.p_2"},
      {h4 ? h4->p_3 : nullptr, "hydro4d.p_3"},     {h4 ? h4->p_4 : nullptr, "hydro4d.p_4"},
      {h4 ? h4->d_1 : nullptr, "hydro4d.d_1"},     {h4 ? h4->d_2 : nullptr, "hydro4d.d_2"},
      {h4 ? h4->nbin : nullptr, "hydro4d.nbin"},
      {h4 ? h4->fs_rho : nullptr, "hydro4d.fs_rho"},
      {h4 ? h4->fs_d : nullptr, "hydro4d.fs_d"},
      {h4 ? h4->fs_d_bound : nullptr, "hydro4d.fs_d_bound"},
      {h4 ? h4->fs_n : nullptr, "hydro4d.fs_n"},
      {h4 ? h4->fs_mass : nullptr, "hydro4d.fs_mass"},
      {h4 ? h4->fs_as_ratio : nullptr, "hydro4d.fs_as_ratio"},
      {h4 ? h4->fs_canting : nullptr, "hydro4d.fs_canting"},
      {h4 ? h4->fs_rg_kappa : nullptr, "hydro4d.fs_rg_kappa"},
      {h4 ? h4->fs_rg_beta : nullptr, "hydro4d.fs_rg_beta"},
      {h4 ? h4->fs_rg_gamma : nullptr, "hydro4d.fs_rg_gamma"},
      {h4 ? h4->fs_rg_zeta : nullptr, "hydro4d.fs_rg_zeta"},
      {r->tb, "results.tb"},                 {r->emissivity, "results.emissivity"},
      {r->angles_deg, "results.angles_deg"}, {r->status, "results.status"},
      {r->kextatmo, "results.kextatmo"},     {r->kexttot, "results.kexttot"},
      {r->salbedo, "results.salbedo"},       {r->nlegen, "results.nlegen"},
      {r->ndoubles, "results.ndoubles"},     {r->ze, "results.ze"},
      {r->att_hydro, "results.att_hydro"},   {r->att_atmo, "results.att_atmo"},
      {r->radar_hgt, "results.radar_hgt"},
  };
  for (const auto &a : arrays)
    if (a.ptr && !device_accessible(a.ptr)) return a.name;
  return nullptr;
}

extern "C" int pam_run_device(pam_context *c, const pam_settings *st, const pam_hydro_desc *hyd,
                              const pam_hydro_4d *h4, const pam_profile *p, const double *freqs,
                              int32_t nfreq, pam_results *res, void *stream) {
  if (c && p && res) {
    if (hipSetDevice(c->device) != hipSuccess) return PAM_FATAL;
    if (const char *bad = first_host_array(h4, p, res)) {
      char m[160];
      snprintf(m, sizeof(m), "pam_run_device: %s is not device memory", bad);
      set_msg(res, m);
      return PAM_FATAL;
    }
  }
  return run_device(c, st, hyd, h4, p, freqs, nfreq, res, (hipStream_t)stream);
}

extern "C" int pam_run_device_noalloc(pam_context *c, const pam_settings *st,
                                      const pam_hydro_desc *hyd, const pam_hydro_4d *h4,
                                      const pam_profile *p, const double *freqs, int32_t nfreq,
                                      pam_results *res, void *stream) {
  if (c && p && res) {
    if (hipSetDevice(c->device) != hipSuccess) return PAM_FATAL;
    if (const char *bad = first_host_array(h4, p, res)) {
      char m[160];
      snprintf(m, sizeof(m), "pam_run_device_noalloc: %s is not device memory", bad);
      set_msg(res, m);
      return PAM_FATAL;
    }
  }
  if (!c || !c->scratch) {
    set_msg(res, "pam_run_device_noalloc: no context or no scratch on its device");
    return PAM_FATAL;
  }
  Arena a;
  a.base = (char *)c->scratch;
  a.size = PAM_NOALLOC_SCRATCH_BYTES;
  return run_device(c, st, hyd, h4, p, freqs, nfreq, res, (hipStream_t)stream, &a);
}

// Host arrays in, host arrays out: stage through one device allocation.
extern "C" int pam_run(pam_context *c, const pam_settings *st, const pam_hydro_desc *hyd,
                       const pam_hydro_4d *h4, const pam_profile *p, const double *freqs,
                       int32_t nfreq, pam_results *res) {
  if (!c || !p || !res) return PAM_FATAL;
  if (hipSetDevice(c->device) != hipSuccess) return PAM_FATAL;
  const size_t nx = p->ngridx, ny = p->ngridy, ncol = nx * ny, ML = p->max_nlyrs,
               nh = p->nhydro, nout = p->noutlevels, nf = nfreq;
  struct Arr {
    const void *h;
    size_t bytes;
    const void **slot;
  };
  pam_profile dp = *p;
  pam_hydro_4d d4;
  memset(&d4, 0, sizeof(d4));
  if (h4) d4 = *h4;
  std::vector<Arr> in;
  auto reg = [&](const void *const *field, size_t bytes) {
    if (*field) in.push_back({*field, bytes, (const void **)field});
  };
  reg((const void *const *)&dp.nlyrs, ncol * 4);
  reg((const void *const *)&dp.hgt_lev, ncol * (ML + 1) * 8);
  reg((const void *const *)&dp.temp_lev, ncol * (ML + 1) * 8);
  reg((const void *const *)&dp.press_lev, ncol * (ML + 1) * 8);
  reg((const void *const *)&dp.relhum_lev, ncol * (ML + 1) * 8);
  reg((const void *const *)&dp.hgt, ncol * ML * 8);
  reg((const void *const *)&dp.temp, ncol * ML * 8);
  reg((const void *const *)&dp.press, ncol * ML * 8);
  reg((const void *const *)&dp.relhum, ncol * ML * 8);
  reg((const void *const *)&dp.hydro_q, ncol * ML * nh * 8);
  reg((const void *const *)&dp.hydro_n, ncol * ML * nh * 8);
  reg((const void *const *)&dp.hydro_reff, ncol * ML * nh * 8);
  reg((const void *const *)&dp.groundtemp, ncol * 8);
  reg((const void *const *)&dp.wind10u, ncol * 8);
  reg((const void *const *)&dp.wind10v, ncol * 8);
  reg((const void *const *)&dp.sfc_salinity, ncol * 8);
  reg((const void *const *)&dp.sfc_emissivity, ncol * 8);
  reg((const void *const *)&dp.sfc_type, ncol * 4);
  reg((const void *const *)&dp.sfc_model, ncol * 4);
  reg((const void *const *)&dp.sfc_refl, ncol);
  reg((const void *const *)&dp.obs_height, ncol * nout * 8);
  const size_t n4 = ncol * ML * nh;
  reg((const void *const *)&d4.as_ratio, n4 * 8);
  reg((const void *const *)&d4.rho_ms, n4 * 8);
  reg((const void *const *)&d4.a_ms, n4 * 8);
  reg((const void *const *)&d4.b_ms, n4 * 8);
  reg((const void *const *)&d4.alpha_as, n4 * 8);
  reg((const void *const *)&d4.beta_as, n4 * 8);
  reg((const void *const *)&d4.p_1, n4 * 8);
  reg((const void *const *)&d4.p_2, n4 * 8);
  reg((const void *const *)&d4.p_3, n4 * 8);
  reg((const void *const *)&d4.p_4, n4 * 8);
  reg((const void *const *)&d4.d_1, n4 * 8);
  reg((const void *const *)&d4.d_2, n4 * 8);
  reg((const void *const *)&d4.nbin, n4 * 4);
  if (d4.fs_nbin > 0) {
    const size_t n5 = n4 * (size_t)d4.fs_nbin;
    reg((const void *const *)&d4.fs_rho, n5 * 8);
    reg((const void *const *)&d4.fs_d, n5 * 8);
    reg((const void *const *)&d4.fs_d_bound, (n5 + n4) * 8);
    reg((const void *const *)&d4.fs_n, n5 * 8);
    reg((const void *const *)&d4.fs_mass, n5 * 8);
    reg((const void *const *)&d4.fs_as_ratio, n5 * 8);
    reg((const void *const *)&d4.fs_canting, n5 * 8);
    reg((const void *const *)&d4.fs_rg_kappa, n5 * 8);
    reg((const void *const *)&d4.fs_rg_beta, n5 * 8);
    reg((const void *const *)&d4.fs_rg_gamma, n5 * 8);
    reg((const void *const *)&d4.fs_rg_zeta, n5 * 8);
  }
  pam_results dr = *res;
  struct Out {
    void *h;
    size_t bytes;
    void **slot;
  };
  std::vector<Out> out;
  auto oreg = [&](void **field, size_t bytes) {
    if (*field) out.push_back({*field, bytes, field});
  };
  oreg((void **)&dr.tb, ncol * nout * PAM_NANGLES * nf * 2 * 8);
  oreg((void **)&dr.emissivity, ncol * 2 * nf * NMU * 8);
  oreg((void **)&dr.angles_deg, PAM_NANGLES * 8);
  oreg((void **)&dr.status, ncol * nf * 4);
  oreg((void **)&dr.kextatmo, ncol * ML * nf * 8);
  oreg((void **)&dr.kexttot, ncol * ML * nf * 8);
  oreg((void **)&dr.salbedo, ncol * ML * nf * 8);
  oreg((void **)&dr.nlegen, ncol * ML * nf * 4);
  oreg((void **)&dr.ndoubles, ncol * ML * nf * 4);
  if (st && st->active) {
    const size_t npol = (size_t)std::max(1, std::min(5, (int)st->radar_npol));
    oreg((void **)&dr.ze, ncol * ML * nf * npol * 8);
    oreg((void **)&dr.att_hydro, ncol * ML * nf * 8);
    oreg((void **)&dr.att_atmo, ncol * ML * nf * 8);
    oreg((void **)&dr.radar_hgt, ncol * ML * 8);
  } else {
    dr.ze = dr.att_hydro = dr.att_atmo = dr.radar_hgt = nullptr;
  }
  size_t total = 0;
  for (auto &a : in) total += (a.bytes + 255) & ~(size_t)255;
  for (auto &o : out) total += (o.bytes + 255) & ~(size_t)255;
  if (c->iobuf.ensure(total + 4096) != hipSuccess) {
    set_msg(res, "pam_run: device allocation failed");
    return PAM_FATAL;
  }
  // everything on the context's own non-blocking stream, so that contexts
  // driven from different host threads overlap on the device
  Carver cv(c->iobuf.p);
  for (auto &a : in) {
    void *dptr = cv.take<char>(a.bytes);
    if (hipMemcpyAsync(dptr, a.h, a.bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
      set_msg(res, "pam_run: copy to device failed");
      return PAM_FATAL;
    }
    *a.slot = dptr;
  }
  for (auto &o : out) {
    void *dptr = cv.take<char>(o.bytes);
    *o.slot = dptr;
  }
  if (dr.tb) (void)hipMemsetAsync(dr.tb, 0, ncol * nout * PAM_NANGLES * nf * 2 * 8, c->stream);
  int rc = run_device(c, st, hyd, h4 ? &d4 : nullptr, &dp, freqs, nfreq, &dr, c->stream);
  for (auto &o : out) {
    if (hipMemcpyAsync(o.h, *o.slot, o.bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess) {
      set_msg(res, "pam_run: copy from device failed");
      return PAM_FATAL;
    }
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    set_msg(res, "pam_run: stream synchronisation failed");
    return PAM_FATAL;
  }
  return rc;
}
