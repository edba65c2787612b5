// y3g_runtime.hip -- host side of liby3g.so: the C ABI of include/y3g.h.
//
// One y3g_runtime per process and device (DeviceInitializer of
// src/utils/CosmoSISSICUDAModule.hh:59-83 binds device = local rank).  It
// owns the per-point tables, the cubature region slab and the integral /
// result buffers, all resident in HBM across calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "y3g.h"
#include "y3g_internal.hpp"
#include "y3g_single.hpp"
#include "y3g_stream.hpp"

using namespace y3g;

namespace {

thread_local std::string g_err;

int fail(int code, std::string const& msg)
{
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                           \
  do {                                                                      \
    hipError_t const e_ = (x);                                              \
    if (e_ != hipSuccess)                                                   \
      return fail(Y3G_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T> int dalloc(T** p, size_t n)
{
  if (n == 0) n = 1;
  hipError_t const e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
  if (e != hipSuccess)
    return fail(Y3G_ERR_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  return 0;
}

template <typename T> void dfree(T*& p)
{
  if (p) (void)hipFree(p);
  p = nullptr;
}

// n Gauss-Legendre nodes and weights on [a, b]: Newton iteration on the
// Legendre polynomial P_n from the Chebyshev-like first guess
// cos(pi (i + 3/4) / (n + 1/2)), weights 2 / ((1 - x^2) P_n'(x)^2).
void gl_nodes(int n, double a, double b, double* x, double* w)
{
  double const c = 0.5 * (a + b), h = 0.5 * (b - a);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double t = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        double const p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p1 = t, p0 = 1.0;
      dp = n * (t * p1 - p0) / (t * t - 1.0);
      double const dt = p1 / dp;
      t -= dt;
      if (std::fabs(dt) < 1e-16) break;
    }
    {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        double const p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p1 = t, p0 = 1.0;
      dp = n * (t * p1 - p0) / (t * t - 1.0);
    }
    double const wi = 2.0 / ((1.0 - t * t) * dp * dp);
    x[i] = c - h * t;
    x[n - 1 - i] = c + h * t;
    w[i] = w[n - 1 - i] = h * wi;
  }
}

}  // namespace

// Scratch memory of y3g_integrate_noalloc: the integral / result staging,
// the work queue, the evaluation counter and the cubature region slab of a
// call live here instead of in buffers the runtime allocates.  One array per
// device (each device holds its own copy, at its own address, until the
// process ends).  Host code never uses the name `scratch` as an address: it
// uses the per-device address from hipGetSymbolAddress (scratch_address).
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Device address of `scratch` on `device` (current), looked up once per
// device.
std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_addr;

int scratch_address(int device, unsigned char** out)
{
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_addr.size() <= device) g_scratch_addr.resize(device + 1, nullptr);
  if (!g_scratch_addr[device]) {
    void* p = nullptr;
    hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(Y3G_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                 hipGetErrorString(e));
    g_scratch_addr[device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_addr[device];
  return 0;
}

}  // namespace

struct y3g_runtime {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int n_cu = 0;
  double clock_ghz = 0.0;
  size_t hbm = 0;
  std::string name;
  // tables
  int n_points = 0;
  y3g_table_spec spec{};
  std::vector<PointDev> hpoints;
  PointDev* points = nullptr;
  double *hmf64 = nullptr, *da64 = nullptr, *sci64 = nullptr;
  float *hmf32 = nullptr, *da32 = nullptr, *sci32 = nullptr;
  int hmf_stride = 0, da_stride = 0, sci_stride = 0;
  double* hmfc64 = nullptr;  // the HMF tables' spline coefficients
  float* hmfc32 = nullptr;
  int hmfc_stride = 0;
  // cubature slab
  double *slab_lo = nullptr, *slab_hi = nullptr, *slab_est = nullptr,
         *slab_err = nullptr, *slab_raw = nullptr;
  int8_t* slab_ax = nullptr;
  int slab_slots = 0, slab_cap = 0;
  unsigned slab_gen = 0;  // bumped when the slab is reallocated (graphs hold its address)
  // batch buffers
  y3g_integral* d_igs = nullptr;
  y3g_result* d_res = nullptr;
  int* d_order = nullptr;
  size_t batch_cap = 0;
  int* d_queue = nullptr;
  unsigned long long* d_nevals = nullptr;
  unsigned char* scratch = nullptr;  // this device's scratch (y3g_integrate_noalloc)
  double last_ms = 0.0;
  long long last_nevals = 0;
  // plain path workspace
  SingleArgs sg{};
  int sg_cap = 0;
  y3g_integral* d_one = nullptr;
  // fixed-GL counts: selection functions per point (host copies, uploaded
  // as one buffer when changed), nodes and results
  struct HostSel {
    std::vector<double> lnm, z, s;
    int n_bins = 0;
  };
  std::vector<HostSel> sel;
  bool sel_dirty = true;
  double* d_seldata = nullptr;
  SelPoint* d_selpts = nullptr;
  double* d_selnodes = nullptr;
  size_t selnodes_cap = 0, seldata_cap = 0, selpts_cap = 0;
  y3g_selgl_spec selnodes_spec{};  // the spec whose nodes d_selnodes holds
  bool selnodes_valid = false;
  double* d_selout = nullptr;
  size_t selout_cap = 0;
  // DES data per point (Y3G_DSIGMA6D_DES), uploaded as one buffer when changed
  std::vector<std::vector<double>> des;
  bool des_dirty = true;
  double* d_des = nullptr;
  long long* d_des_off = nullptr;
  size_t des_cap = 0, des_off_cap = 0;
  // NFW_DSIGMA_MIS table {ln x [nx], ln xmis [nm], T [nm][nx]}
  double* d_nfwmis = nullptr;
  int nfw_nx = 0, nfw_nm = 0;
};

// SIG_MAX tables packed as sig_max() reads them: sizes, increasing axes.
static bool sigmax_ok(const double* par, long long npar)
{
  if (npar < 3) return false;
  double const a = par[0], b = par[1], c = par[2];
  if (!(a >= 2.0 && b >= 2.0 && c >= 2.0 && a <= 4096.0 && b <= 4096.0 && c <= 4096.0) ||
      a != double(int(a)) || b != double(int(b)) || c != double(int(c)))
    return false;
  long long const nr = (long long)a, nm = (long long)b, nz = (long long)c;
  if (npar != 3 + nr + nm + nz + nm * nr + nz * nr + nm * nz) return false;
  long long const n[3] = {nr, nm, nz};
  long long off = 3;
  for (int ax = 0; ax < 3; off += n[ax], ++ax)
    for (long long k = 1; k < n[ax]; ++k)
      if (!(par[off + k] > par[off + k - 1])) return false;
  return true;
}

// The points' DES data on the device (one buffer, offsets per point).
static int ensure_des(y3g_runtime* rt)
{
  if (!rt->des_dirty) return 0;
  // offsets for every point with tables (the kernels read des_off[pt] for
  // any point of a launch) plus one past the end: point p's data is
  // [off[p], off[p + 1]), empty for points without DES data
  std::vector<double> data;
  size_t const np = std::max<size_t>(rt->des.size(), (size_t)std::max(rt->n_points, 0));
  std::vector<long long> off(np + 1, 0);
  for (size_t p = 0; p < np; ++p) {
    off[p] = (long long)data.size();
    if (p < rt->des.size()) data.insert(data.end(), rt->des[p].begin(), rt->des[p].end());
  }
  off[np] = (long long)data.size();
  if (data.empty()) data.push_back(0.0);
  if (data.size() > rt->des_cap) {
    dfree(rt->d_des);
    rt->des_cap = 0;
    if (int rc = dalloc(&rt->d_des, data.size())) return rc;
    rt->des_cap = data.size();
  }
  if (off.size() > rt->des_off_cap) {
    dfree(rt->d_des_off);
    rt->des_off_cap = 0;
    if (int rc = dalloc(&rt->d_des_off, off.size())) return rc;
    rt->des_off_cap = off.size();
  }
  HIPCHK(hipMemcpy(rt->d_des, data.data(), sizeof(double) * data.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(rt->d_des_off, off.data(), sizeof(long long) * off.size(),
                   hipMemcpyHostToDevice));
  rt->des_dirty = false;
  return 0;
}

// Dimensions of a real-data integral: the descriptor's (the miscentred kinds
// integrate lambda_obs in closed form) and the reference volume's.
static int des_ndim(int kind)
{
  return kind == Y3G_DSIGMA6D_DES ? 5 : kind == Y3G_COUNT5D_DES ? 4
       : kind == Y3G_COUNT4D_CARD ? 4 : 3;
}
static int des_volume_ndim(int kind)
{
  return kind == Y3G_DSIGMA6D_DES ? 6 : kind == Y3G_COUNT5D_DES ? 5 : des_ndim(kind);
}

// The P[X] operator's point data packed as integrand_pop_k() reads them:
// {n_chi, n_bm, n_bz, n_xr, n_xz, chi_z, chi_dc, bias_lnM, bias_z, bias,
// xi_r, xi_z, xi}, each axis increasing.
static bool pop_ok(const double* par, long long npar)
{
  if (npar < 5) return false;
  long long n[5];
  for (int k = 0; k < 5; ++k) {
    if (!(par[k] >= 2.0 && par[k] <= 65536.0) || par[k] != double((long long)par[k]))
      return false;
    n[k] = (long long)par[k];
  }
  long long const need = 5 + 2 * n[0] + n[1] + n[2] + n[1] * n[2] + n[3] + n[4] + n[3] * n[4];
  if (npar != need) return false;
  auto inc = [&](long long off, long long m) {
    for (long long k = 1; k < m; ++k)
      if (!(par[off + k] > par[off + k - 1])) return false;
    return true;
  };
  long long o = 5;
  if (!inc(o, n[0])) return false;                       // chi_z
  o += 2 * n[0];
  if (!inc(o, n[1]) || !inc(o + n[1], n[2])) return false;  // bias lnM, z
  o += n[1] + n[2] + n[1] * n[2];
  if (!inc(o, n[3]) || !inc(o + n[3], n[4])) return false;  // xi r, z
  return true;
}

// INT_LC_LT_DES_t2 tables packed as int_lc_lt_des() reads them.
static bool lclt_ok(const double* par, long long npar)
{
  if (npar < 2) return false;
  double const a = par[0], b = par[1];
  if (!(a >= 2.0 && b >= 2.0 && a <= 4096.0 && b <= 4096.0) || a != double(int(a)) ||
      b != double(int(b)) || npar != 2 + (long long)(a + b) + 4LL * (long long)(a * b))
    return false;
  for (int k = 1; k < int(a); ++k)
    if (!(par[2 + k] > par[1 + k])) return false;
  for (int k = 1; k < int(b); ++k)
    if (!(par[2 + int(a) + k] > par[1 + int(a) + k])) return false;
  return true;
}

static bool has_des(const y3g_runtime* rt, int point)
{
  return point >= 0 && (size_t)point < rt->des.size() && !rt->des[point].empty();
}

extern "C" {

const char* y3g_last_error(void) { return g_err.c_str(); }
const char* y3g_version(void) { return "y3g 0.1 gfx950"; }

void y3g_options_default(y3g_options* o)
{
  o->precision = Y3G_FP32SPLIT;
  o->max_regions = 262144;
  o->max_split_per_step = 1024;
  o->init_split = 0;
  o->block_threads = 512;
  o->slots = 0;
  o->vegas_calls = 1000000;
  o->vegas_itmax = 5;
  o->vegas_itskip = 1;
  o->vegas_nbins = 64;
  o->vegas_alpha = 1.5;
  o->single = 0;
  o->err_c = 0.0;
  o->err_smin = 1.0;
  o->n_bands = 0;
  for (int k = 0; k < 8; ++k) o->band_eps[k] = 0.0, o->band_regions[k] = 0;
  o->tl_c = 0.0;
  o->tl_smin = 1.0 / 16.0;
  o->hmf_smooth = 1;
  o->init_split_lnm = 4;
  o->err_scale = 1.5;
  o->ext_ratio = 0.0;
  o->ext_regions = 0;
}

void y3g_default_params(y3g_params* p)
{
  p->omega_m = 0.3;
  p->h = 0.7;
  p->hmf_amp = 0.8;
  p->hmf_A = 0.2;
  p->hmf_a = 1.5;
  p->hmf_c = 1.2;
  p->mor_lnl0 = 3.2;
  p->mor_slope = 0.9;
  p->mor_zslope = 0.5;
  p->mor_sigma_ln = 0.25;
  p->f_prj = 0.2;
  p->prj_scale = 0.1;
  p->f_mis = 0.2;
  p->mis_rlambda = 0.2;
  p->z_source = 1.2;
  p->sigz0 = 0.01;
  p->hmf_s = 0.0;
  p->hmf_q = 1.0;
}

void y3g_spec_config0(y3g_table_spec* s)
{
  s->lnm_lo = 29.933606208922594;  // ln 1e13
  s->lnm_hi = 35.690068941407709;  // 15.5 ln 10
  s->n_lnm = 200;
  s->z_lo = 0.20;
  s->z_hi = 0.35;
  s->n_z = 32;
  s->da_zmax = 1.0;
  s->n_da = 512;
  s->sci_zmax = 1.0;
  s->n_sci = 128;
}

void y3g_spec_wide(y3g_table_spec* s)
{
  y3g_spec_config0(s);
  s->z_lo = 0.10;
  s->z_hi = 0.75;
}

void y3g_make_integral(const y3g_params* p, const y3g_table_spec* s, int kind,
                       double lo_lo, double lo_hi, double zo_lo, double zo_hi,
                       double radius, y3g_integral* ig)
{
  std::memset(ig, 0, sizeof(*ig));
  ig->kind = kind;
  ig->lo_lo = lo_lo;
  ig->lo_hi = lo_hi;
  ig->zo_lo = zo_lo;
  ig->zo_hi = zo_hi;
  ig->radius = radius;
  ig->lnlt_lo = 0.0;                   // ln 1
  ig->lnlt_hi = 5.7037824746562009;    // ln 300
  ig->eps_rel = 1e-4;
  ig->eps_abs = 1e-300;
  ig->max_eval = 1LL << 40;
  ig->seed = 0x59334331ull;
  double const zt_lo = kind == Y3G_COUNT3D ? s->z_lo : zo_lo - 0.1;
  double const zt_hi = kind == Y3G_COUNT3D ? s->z_hi : zo_hi + 0.1;
  switch (kind) {
    case Y3G_COUNT3D:
      ig->ndim = 3;
      ig->lows[0] = 0.0, ig->highs[0] = 1.0;
      ig->lows[1] = zt_lo, ig->highs[1] = zt_hi;
      ig->lows[2] = s->lnm_lo, ig->highs[2] = s->lnm_hi;
      break;
    case Y3G_COUNT5D:
      ig->ndim = 5;
      ig->lows[0] = lo_lo, ig->highs[0] = lo_hi;
      ig->lows[1] = 0.0, ig->highs[1] = 1.0;  // standardised photo-z (zo_std)
      ig->lows[2] = 0.0, ig->highs[2] = 1.0;
      ig->lows[3] = zt_lo, ig->highs[3] = zt_hi;
      ig->lows[4] = s->lnm_lo, ig->highs[4] = s->lnm_hi;
      break;
    default: {
      ig->ndim = 6;
      double const rl = std::pow(0.5 * (lo_lo + lo_hi) / 100.0, 0.2);
      ig->sigma_mis = p->mis_rlambda * rl;
      ig->rmis_norm = 1.0 / (1.0 - std::exp(-0.5 * 6.0 * 6.0));
      ig->lows[0] = lo_lo, ig->highs[0] = lo_hi;
      ig->lows[1] = 0.0, ig->highs[1] = 1.0;
      ig->lows[2] = zt_lo, ig->highs[2] = zt_hi;
      ig->lows[3] = s->lnm_lo, ig->highs[3] = s->lnm_hi;
      ig->lows[4] = 0.0, ig->highs[4] = -std::expm1(-0.5 * 6.0 * 6.0);
      ig->lows[5] = 0.0, ig->highs[5] = 3.14159265358979323846;
      break;
    }
  }
}

int y3g_make_integral_volume(const y3g_params* p, const y3g_table_spec* s,
                             int kind, const double* lows, const double* highs,
                             const double* grid, int n_grid, y3g_integral* ig)
{
  if (!p || !s || !lows || !highs || !ig)
    return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: null argument");
  if (is_pop_kind(kind)) {
    // P_operator_gpu: volume (lt, zt, lnm), grid (zo_low, zo_high,
    // lambda_bin) (p_operator_gpu_t.cuh:172-179, 278-321)
    if (n_grid != 3 || !grid)
      return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: the P[X] kinds take (zo_low, "
                               "zo_high, lambda_bin)");
    for (int i = 0; i < 3; ++i)
      if (!(lows[i] <= highs[i]) || !std::isfinite(lows[i]) || !std::isfinite(highs[i]))
        return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: axis " + std::to_string(i) +
                                     " has low > high or a non-finite bound");
    if (!(lows[1] >= 0.0) || !(highs[1] <= s->da_zmax))
      return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: zt range outside the distance "
                               "table [0, " + std::to_string(s->da_zmax) + "]");
    if (!(grid[0] <= grid[1]) || !(grid[2] >= 0.0 && grid[2] <= 3.0) ||
        grid[2] != double(int(grid[2])))
      return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: grid needs zo_low <= zo_high and "
                               "lambda_bin in 0..3");
    std::memset(ig, 0, sizeof(*ig));
    ig->kind = kind;
    ig->ndim = 3;
    ig->zo_lo = grid[0];
    ig->zo_hi = grid[1];
    ig->lo_lo = grid[2];
    ig->lo_hi = pop_lob_center(int(grid[2]));
    for (int i = 0; i < 3; ++i) ig->lows[i] = lows[i], ig->highs[i] = highs[i];
    ig->eps_rel = 1e-4;
    ig->eps_abs = 1e-300;
    ig->max_eval = 1LL << 40;
    ig->seed = 0x59334331ull;
    return 0;
  }
  if (is_des_kind(kind)) {
    // the reference's volume and its grid point: (zo_low, zo_high, radius)
    // for Delta-Sigma, (zo_low, zo_high) for the counts; the miscentred
    // kinds' lo axis becomes the closed-form lambda_obs bin (lo_lo, lo_hi)
    int const nd = des_volume_ndim(kind);
    bool const sig = kind == Y3G_DSIGMA6D_DES || kind == Y3G_DSIGMA3D_DES;
    int const ng = sig ? 3 : 2;
    if (n_grid != ng || !grid)
      return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: kind " + std::to_string(kind) +
                                   " takes " + std::to_string(ng) + " grid values, got " +
                                   std::to_string(n_grid));
    for (int i = 0; i < nd; ++i)
      if (!(lows[i] <= highs[i]) || !std::isfinite(lows[i]) || !std::isfinite(highs[i]))
        return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: axis " + std::to_string(i) +
                                     " has low > high or a non-finite bound");
    bool const miscent = kind == Y3G_DSIGMA6D_DES || kind == Y3G_COUNT5D_DES;
    int const iz = (miscent || kind == Y3G_COUNT4D_CARD) ? 2 : 1;
    if (miscent && !(lows[1] > 0.0))
      return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: lc range must be > 0");
    if (!(lows[iz] >= 0.0) || !(highs[iz] <= s->da_zmax))
      return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: zt range outside the distance "
                               "table [0, " + std::to_string(s->da_zmax) + "]");
    if (!(grid[0] <= grid[1]) || (sig && !(grid[2] > 0.0)))
      return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: grid needs zo_low <= zo_high "
                               "(and radius > 0)");
    std::memset(ig, 0, sizeof(*ig));
    ig->kind = kind;
    ig->ndim = des_ndim(kind);
    ig->zo_lo = grid[0];
    ig->zo_hi = grid[1];
    ig->radius = sig ? grid[2] : 0.0;
    int const skip = miscent ? 1 : 0;
    if (miscent) ig->lo_lo = lows[0], ig->lo_hi = highs[0];
    for (int i = skip; i < nd; ++i) ig->lows[i - skip] = lows[i], ig->highs[i - skip] = highs[i];
    ig->eps_rel = 1e-4;
    ig->eps_abs = 1e-300;
    ig->max_eval = 1LL << 40;
    ig->seed = 0x59334331ull;
    return 0;
  }
  int const nd = kind == Y3G_COUNT3D ? 3 : kind == Y3G_COUNT5D ? 5
               : kind == Y3G_DSIGMA6D ? 6 : 0;
  int const ng = kind == Y3G_COUNT3D ? 2 : kind == Y3G_COUNT5D ? 0 : 3;
  if (nd == 0) return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: unknown kind");
  if (n_grid != ng || (ng > 0 && !grid))
    return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: kind " + std::to_string(kind) +
                                 " takes " + std::to_string(ng) + " grid values, got " +
                                 std::to_string(n_grid));
  for (int i = 0; i < nd; ++i)
    if (!(lows[i] <= highs[i]) || !std::isfinite(lows[i]))
      return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: axis " + std::to_string(i) +
                                   " has low > high or a non-finite low");
  // physical axis positions of (lc, zt, lnm) per kind
  int const ilc = kind == Y3G_COUNT3D ? 0 : kind == Y3G_COUNT5D ? 2 : 1;
  int const izt = ilc + 1, ilm = ilc + 2;
  double const zmax = kind == Y3G_DSIGMA6D ? std::min(s->da_zmax, s->sci_zmax) : s->da_zmax;
  if (!(lows[ilc] > 0.0) || !std::isfinite(highs[ilc]))
    return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: lc range must be finite and > 0");
  if (!(lows[izt] >= 0.0) || !(highs[izt] <= zmax))
    return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: zt range outside the distance "
                             "tables [0, " + std::to_string(zmax) + "]");
  if (!std::isfinite(highs[ilm]))
    return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: lnm range must be finite");
  if (kind == Y3G_COUNT3D) {
    y3g_make_integral(p, s, kind, grid[0], grid[1], 0.0, 0.0, 0.0, ig);
  } else if (kind == Y3G_COUNT5D) {
    y3g_make_integral(p, s, kind, lows[0], highs[0], lows[1], highs[1], 0.0, ig);
  } else {
    if (!(grid[2] > 0.0) || !(grid[0] <= grid[1]))
      return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: grid needs zo_low <= zo_high "
                               "and radius > 0");
    y3g_make_integral(p, s, kind, lows[0], highs[0], grid[0], grid[1], grid[2], ig);
  }
  ig->lnlt_lo = std::log(lows[ilc]);
  ig->lnlt_hi = std::log(highs[ilc]);
  ig->lows[izt] = lows[izt], ig->highs[izt] = highs[izt];
  ig->lows[ilm] = lows[ilm], ig->highs[ilm] = highs[ilm];
  if (kind == Y3G_DSIGMA6D) {
    // offset radius as its Rayleigh CDF u = 1 - exp(-r^2 / 2 sigma^2), no
    // renormalisation (the pdf is integrated over [rmis_low, rmis_high]),
    // cut at u = 1 - exp(-18) (6 sigma); theta over [a, b] with the 1/2pi
    // of the angular average, folded onto [0, pi] for the full circle
    double const sg = ig->sigma_mis;
    if (!(lows[4] >= 0.0))
      return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: rmis_low must be >= 0");
    double const twopi = 6.283185307179586;
    if (!(lows[5] >= 0.0) || !(highs[5] <= twopi * (1.0 + 1e-12)))
      return fail(Y3G_ERR_ARG, "y3g_make_integral_volume: theta range must lie in [0, 2 pi]");
    double const ucut = -std::expm1(-18.0);
    auto cdf = [sg](double r) {
      return std::isfinite(r) ? -std::expm1(-0.5 * (r / sg) * (r / sg)) : 1.0;
    };
    ig->lows[4] = std::min(cdf(lows[4]), ucut);
    ig->highs[4] = std::min(cdf(highs[4]), ucut);
    if (lows[5] == 0.0 && std::fabs(highs[5] - twopi) <= 1e-12 * twopi) {
      ig->lows[5] = 0.0, ig->highs[5] = 3.14159265358979323846;
      ig->rmis_norm = 1.0;
    } else {
      ig->lows[5] = lows[5], ig->highs[5] = highs[5];
      ig->rmis_norm = 0.5;
    }
  }
  return Y3G_OK;
}

int y3g_runtime_create(int device, y3g_runtime** out)
{
  if (!out) return fail(Y3G_ERR_ARG, "y3g_runtime_create: null out");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return fail(Y3G_ERR_NODEVICE, "no HIP device");
  if (device < 0 || device >= ndev)
    return fail(Y3G_ERR_ARG, "device " + std::to_string(device) + " of " +
                               std::to_string(ndev));
  HIPCHK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(Y3G_ERR_NODEVICE,
                std::string("liby3g is built for gfx950, device is ") +
                  prop.gcnArchName);
  auto* rt = new y3g_runtime;
  rt->device = device;
  rt->n_cu = prop.multiProcessorCount;
  rt->clock_ghz = prop.clockRate / 1e6;
  rt->hbm = prop.totalGlobalMem;
  rt->name = prop.gcnArchName;
  if (hipStreamCreateWithFlags(&rt->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&rt->ev0) != hipSuccess ||
      hipEventCreate(&rt->ev1) != hipSuccess ||
      dalloc(&rt->d_queue, 1) || dalloc(&rt->d_nevals, 1)) {
    y3g_runtime_destroy(rt);
    return fail(Y3G_ERR_HIP, "runtime resources");
  }
  if (int rc = scratch_address(device, &rt->scratch)) {
    std::string const msg = g_err;
    y3g_runtime_destroy(rt);
    return fail(rc, msg);
  }
  *out = rt;
  return 0;
}

int y3g_runtime_destroy(y3g_runtime* rt)
{
  if (!rt) return 0;
  (void)hipSetDevice(rt->device);
  if (rt->stream) (void)hipStreamSynchronize(rt->stream);
  dfree(rt->points);
  dfree(rt->hmf64), dfree(rt->da64), dfree(rt->sci64);
  dfree(rt->hmf32), dfree(rt->da32), dfree(rt->sci32);
  dfree(rt->hmfc64), dfree(rt->hmfc32);
  dfree(rt->slab_lo), dfree(rt->slab_hi), dfree(rt->slab_est),
    dfree(rt->slab_err), dfree(rt->slab_raw), dfree(rt->slab_ax);
  dfree(rt->d_igs), dfree(rt->d_res), dfree(rt->d_order);
  dfree(rt->d_nfwmis);
  dfree(rt->d_queue), dfree(rt->d_nevals);
  dfree(rt->sg.lo), dfree(rt->sg.hi), dfree(rt->sg.est), dfree(rt->sg.err);
  dfree(rt->sg.ax), dfree(rt->sg.list), dfree(rt->sg.sel), dfree(rt->sg.pest);
  dfree(rt->sg.perr), dfree(rt->sg.partial), dfree(rt->sg.IE), dfree(rt->sg.hc);
  dfree(rt->sg.hq), dfree(rt->sg.blockcnt), dfree(rt->sg.blockoff), dfree(rt->d_one);
  dfree(rt->d_seldata), dfree(rt->d_selpts), dfree(rt->d_selnodes), dfree(rt->d_selout);
  dfree(rt->d_des), dfree(rt->d_des_off);
  if (rt->ev0) (void)hipEventDestroy(rt->ev0);
  if (rt->ev1) (void)hipEventDestroy(rt->ev1);
  if (rt->stream) (void)hipStreamDestroy(rt->stream);
  delete rt;
  return 0;
}

int y3g_device_info(y3g_runtime* rt, int* n_cu, double* clock_ghz,
                    size_t* hbm_bytes, char* name, int name_len)
{
  if (!rt) return fail(Y3G_ERR_ARG, "null runtime");
  if (n_cu) *n_cu = rt->n_cu;
  if (clock_ghz) *clock_ghz = rt->clock_ghz;
  if (hbm_bytes) *hbm_bytes = rt->hbm;
  if (name && name_len > 0) {
    std::strncpy(name, rt->name.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  return 0;
}

static PointDev make_point(const y3g_params& p, const y3g_table_spec& s)
{
  PointDev P;
  std::memset(&P, 0, sizeof(P));
  P.p = p;
  P.lnm0 = s.lnm_lo;
  P.dlnm = (s.lnm_hi - s.lnm_lo) / (s.n_lnm - 1);
  P.inv_dlnm = 1.0 / P.dlnm;
  P.z0 = s.z_lo;
  P.dz = (s.z_hi - s.z_lo) / (s.n_z - 1);
  P.inv_dz = 1.0 / P.dz;
  P.inv_dzda = (s.n_da - 1) / s.da_zmax;
  P.inv_dzsci = (s.n_sci - 1) / s.sci_zmax;
  P.rho_m = p.omega_m * kRhoCrit;
  P.nm = s.n_lnm;
  P.nz = s.n_z;
  P.nda = s.n_da;
  P.nsci = s.n_sci;
  return P;
}

static int check_spec(const y3g_table_spec* s)
{
  if (s->n_lnm < 2 || s->n_z < 2 || s->n_lnm * s->n_z > kMaxHmf ||
      s->n_da < 2 || s->n_da > kMaxDa || s->n_sci < 2 || s->n_sci > kMaxSci ||
      !(s->lnm_hi > s->lnm_lo) || !(s->z_hi > s->z_lo) || !(s->da_zmax > 0) ||
      !(s->sci_zmax > 0))
    return fail(Y3G_ERR_ARG, "table spec out of range (n_lnm*n_z <= 6400, "
                             "n_da <= 512, n_sci <= 128)");
  return 0;
}

// (n_z + 4) x (n_lnm + 4): the HMF spline coefficients of one point
static int hmfc_stride_of(const y3g_table_spec& s) { return (s.n_lnm + 4) * (s.n_z + 4); }
// tables this small on an axis are read bilinearly (no coefficients)
static bool smooth_ok(const y3g_table_spec& s)
{
  return s.n_lnm >= kSmoothMinKnots && s.n_z >= kSmoothMinKnots;
}

static int alloc_tables(y3g_runtime* rt, const y3g_table_spec* spec, int n)
{
  dfree(rt->points);
  dfree(rt->hmf64), dfree(rt->da64), dfree(rt->sci64);
  dfree(rt->hmf32), dfree(rt->da32), dfree(rt->sci32);
  dfree(rt->hmfc64), dfree(rt->hmfc32);
  rt->n_points = 0;
  rt->spec = *spec;
  rt->hmf_stride = spec->n_lnm * spec->n_z;
  rt->hmfc_stride = hmfc_stride_of(*spec);
  rt->da_stride = spec->n_da;
  rt->sci_stride = spec->n_sci;
  if (dalloc(&rt->points, n) || dalloc(&rt->hmf64, (size_t)n * rt->hmf_stride) ||
      dalloc(&rt->hmf32, (size_t)n * rt->hmf_stride) ||
      dalloc(&rt->da64, (size_t)n * rt->da_stride) ||
      dalloc(&rt->da32, (size_t)n * rt->da_stride) ||
      dalloc(&rt->sci64, (size_t)n * rt->sci_stride) ||
      dalloc(&rt->sci32, (size_t)n * rt->sci_stride) ||
      dalloc(&rt->hmfc64, (size_t)n * rt->hmfc_stride) ||
      dalloc(&rt->hmfc32, (size_t)n * rt->hmfc_stride))
    return Y3G_ERR_MEMORY;
  rt->n_points = n;
  rt->des_dirty = true;  // the DES offsets cover every point with tables
  return 0;
}

int y3g_tables_build(y3g_runtime* rt, const y3g_params* params,
                     const y3g_table_spec* spec, int n_points)
{
  if (!rt || !params || !spec || n_points <= 0)
    return fail(Y3G_ERR_ARG, "y3g_tables_build: bad arguments");
  if (int rc = check_spec(spec)) return rc;
  HIPCHK(hipSetDevice(rt->device));
  if (rt->n_points != n_points || std::memcmp(&rt->spec, spec, sizeof(*spec))) {
    if (int rc = alloc_tables(rt, spec, n_points)) return rc;
  }
  rt->hpoints.resize(n_points);
  for (int k = 0; k < n_points; ++k) rt->hpoints[k] = make_point(params[k], *spec);
  HIPCHK(hipMemcpyAsync(rt->points, rt->hpoints.data(),
                        sizeof(PointDev) * n_points, hipMemcpyHostToDevice,
                        rt->stream));
  TableArgs A;
  A.points = rt->points;
  A.spec = *spec;
  A.hmf64 = rt->hmf64, A.hmf32 = rt->hmf32, A.hmf_stride = rt->hmf_stride;
  A.da64 = rt->da64, A.da32 = rt->da32, A.da_stride = rt->da_stride;
  A.sci64 = rt->sci64, A.sci32 = rt->sci32, A.sci_stride = rt->sci_stride;
  A.hmfc64 = rt->hmfc64, A.hmfc32 = rt->hmfc32, A.hmfc_stride = rt->hmfc_stride;
  HIPCHK(launch_tables(A, n_points, rt->stream));
  HIPCHK(hipStreamSynchronize(rt->stream));
  // the points are new samples: DES data uploaded for the old ones is stale
  rt->des.clear();
  rt->des_dirty = true;
  return 0;
}

int y3g_tables_upload(y3g_runtime* rt, int point, const double* hmf,
                      const double* da, const double* sci)
{
  if (!rt || point < 0 || point >= rt->n_points)
    return fail(Y3G_ERR_ARG, "y3g_tables_upload: no such point");
  HIPCHK(hipSetDevice(rt->device));
  auto up = [&](const double* h, double* d64, float* d32, int n) -> int {
    if (!h) return 0;
    std::vector<float> f(h, h + n);
    HIPCHK(hipMemcpy(d64 + (size_t)point * n, h, sizeof(double) * n,
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d32 + (size_t)point * n, f.data(), sizeof(float) * n,
                     hipMemcpyHostToDevice));
    return 0;
  };
  if (int rc = up(hmf, rt->hmf64, rt->hmf32, rt->hmf_stride)) return rc;
  if (hmf && smooth_ok(rt->spec)) {
    // the uploaded table's spline coefficients (hat-smoothed HMF)
    TableArgs A{};
    A.spec = rt->spec;
    A.hmf64 = rt->hmf64, A.hmf_stride = rt->hmf_stride;
    A.hmfc64 = rt->hmfc64, A.hmfc32 = rt->hmfc32, A.hmfc_stride = rt->hmfc_stride;
    HIPCHK(launch_hmf_coefs(A, point, 1, rt->stream));
    HIPCHK(hipStreamSynchronize(rt->stream));
  }
  if (int rc = up(da, rt->da64, rt->da32, rt->da_stride)) return rc;
  if (int rc = up(sci, rt->sci64, rt->sci32, rt->sci_stride)) return rc;
  // a new sample's tables: the point's earlier DES data is stale
  if ((size_t)point < rt->des.size() && !rt->des[point].empty()) {
    rt->des[point].clear();
    rt->des_dirty = true;
  }
  return 0;
}

int y3g_tables_download(y3g_runtime* rt, int point, double* hmf, double* da,
                        double* sci)
{
  if (!rt || point < 0 || point >= rt->n_points)
    return fail(Y3G_ERR_ARG, "y3g_tables_download: no such point");
  HIPCHK(hipSetDevice(rt->device));
  if (hmf)
    HIPCHK(hipMemcpy(hmf, rt->hmf64 + (size_t)point * rt->hmf_stride,
                     sizeof(double) * rt->hmf_stride, hipMemcpyDeviceToHost));
  if (da)
    HIPCHK(hipMemcpy(da, rt->da64 + (size_t)point * rt->da_stride,
                     sizeof(double) * rt->da_stride, hipMemcpyDeviceToHost));
  if (sci)
    HIPCHK(hipMemcpy(sci, rt->sci64 + (size_t)point * rt->sci_stride,
                     sizeof(double) * rt->sci_stride, hipMemcpyDeviceToHost));
  return 0;
}

static int ensure_batch(y3g_runtime* rt, size_t n)
{
  if (n <= rt->batch_cap) return 0;
  dfree(rt->d_igs), dfree(rt->d_res), dfree(rt->d_order);
  rt->batch_cap = 0;
  size_t const c = std::max<size_t>(n, 1024);
  if (dalloc(&rt->d_igs, c) || dalloc(&rt->d_res, c) || dalloc(&rt->d_order, c))
    return Y3G_ERR_MEMORY;
  rt->batch_cap = c;
  return 0;
}

static int ensure_slab(y3g_runtime* rt, int slots, int cap)
{
  if (slots <= rt->slab_slots && cap == rt->slab_cap) return 0;
  dfree(rt->slab_lo), dfree(rt->slab_hi), dfree(rt->slab_est),
    dfree(rt->slab_err), dfree(rt->slab_raw), dfree(rt->slab_ax);
  rt->slab_slots = rt->slab_cap = 0;
  size_t const n = (size_t)slots * cap;
  if (dalloc(&rt->slab_lo, n * kMaxDim) || dalloc(&rt->slab_hi, n * kMaxDim) ||
      dalloc(&rt->slab_est, n) || dalloc(&rt->slab_err, n) ||
      dalloc(&rt->slab_raw, n) || dalloc(&rt->slab_ax, n + 64))
    return Y3G_ERR_MEMORY;
  rt->slab_slots = slots;
  rt->slab_cap = cap;
  rt->slab_gen += 1;
  return 0;
}

static int ensure_single(y3g_runtime* rt, int cap)
{
  if (cap == rt->sg_cap) return 0;
  SingleArgs& g = rt->sg;
  dfree(g.lo), dfree(g.hi), dfree(g.est), dfree(g.err), dfree(g.ax), dfree(g.list);
  dfree(g.sel), dfree(g.pest), dfree(g.perr), dfree(g.partial), dfree(g.IE);
  dfree(g.hc), dfree(g.hq), dfree(g.blockcnt), dfree(g.blockoff), dfree(rt->d_one);
  rt->sg_cap = 0;
  size_t const nt = (size_t)(cap + 4095) / 4096;
  if (dalloc(&g.lo, (size_t)cap * kMaxDim) || dalloc(&g.hi, (size_t)cap * kMaxDim) ||
      dalloc(&g.est, cap) || dalloc(&g.err, cap) || dalloc(&g.ax, cap) ||
      dalloc(&g.list, 2 * (size_t)cap) || dalloc(&g.sel, cap) || dalloc(&g.pest, cap) ||
      dalloc(&g.perr, cap) || dalloc(&g.partial, 512) || dalloc(&g.IE, 2) ||
      dalloc(&g.hc, 64) || dalloc(&g.hq, 64) || dalloc(&g.blockcnt, 2 * nt) ||
      dalloc(&g.blockoff, 4 * nt) || dalloc(&rt->d_one, 1))
    return Y3G_ERR_MEMORY;
  g.cap = cap;
  rt->sg_cap = cap;
  return 0;
}

// Largest expected cost first: dimension, then the tight tolerances
// (eps_rel < 5e-5: the integrals a budget extension most often takes to
// four times their base budget) before the others, then point by point,
// tighter tolerance first.  The long integrals start early and the short
// ones fill the end of the launch, when workgroups run out of work; keeping
// a point's integrals together within each group keeps the few points the
// resident workgroups are on at a time in L2 (their HMF coefficients: 29 KB
// per point, read by every pair factor).
static void lpt_order(const y3g_integral* igs, const std::vector<int>& ids,
                      std::vector<int>& order)
{
  order = ids;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (igs[a].ndim != igs[b].ndim) return igs[a].ndim > igs[b].ndim;
    bool const ta = igs[a].eps_rel < 5e-5, tb = igs[b].eps_rel < 5e-5;
    if (ta != tb) return ta;
    if (igs[a].point != igs[b].point) return igs[a].point < igs[b].point;
    return igs[a].eps_rel < igs[b].eps_rel;
  });
}

// Workgroups kept resident by the persistent kernels: as many per CU as
// the LDS (160 KB per CU) and 2048 threads per CU allow.
static int resident_slots(const y3g_runtime* rt, const y3g_options& o, int algorithm)
{
  if (o.slots > 0) return o.slots;
  int const block = o.block_threads == 256 ? 256 : 512;
  int per_cu = 2048 / block;
  if (algorithm == Y3G_PAGANI) {
    int const n = cubature_blocks_per_cu(o.precision, block);
    per_cu = n > 0 ? n : 1;
  }
  return rt->n_cu * per_cu;
}

// Device buffers of one y3g_integrate call: the runtime's own (allocated
// and grown on demand), or carved out of the device's scratch array
// (y3g_integrate_noalloc).
struct CallBufs {
  y3g_integral* igs = nullptr;
  y3g_result* res = nullptr;
  int* order = nullptr;
  int* queue = nullptr;
  unsigned long long* nevals = nullptr;
  double *lo = nullptr, *hi = nullptr, *est = nullptr, *err = nullptr, *raw = nullptr;
  int8_t* ax = nullptr;
  int slab_slots = 0;  // workgroups the scratch slab holds at max_regions
  bool noalloc = false;
};

// Zero n words of a call's counters: a fill kernel in the no-allocation
// path (whose callers may capture it), a memset otherwise.
static int zero_words(y3g_runtime* rt, const CallBufs& B, void* p, size_t n)
{
  if (B.noalloc) {
    HIPCHK(launch_zero_words(p, n, rt->stream));
  } else {
    HIPCHK(hipMemsetAsync(p, 0, 4 * n, rt->stream));
  }
  return 0;
}

// Lay out n integrals, their results and order, the counters and (cubature)
// a region slab of max_regions per workgroup in the scratch array: 256-byte
// aligned pieces, as many slab workgroups as the rest holds.
static int carve_scratch(y3g_runtime* rt, size_t n, int cap, bool slab, CallBufs& B)
{
  unsigned char* const base = rt->scratch;
  size_t off = 256;  // queue (int) at 0, evaluation counter (u64) at 8
  auto take = [&](size_t bytes) {
    size_t const at = off;
    off += (bytes + 255) & ~size_t(255);
    return base + at;
  };
  B.noalloc = true;
  B.queue = reinterpret_cast<int*>(base);
  B.nevals = reinterpret_cast<unsigned long long*>(base + 8);
  B.order = reinterpret_cast<int*>(take(n * sizeof(int)));
  B.igs = reinterpret_cast<y3g_integral*>(take(n * sizeof(y3g_integral)));
  B.res = reinterpret_cast<y3g_result*>(take(n * sizeof(y3g_result)));
  std::string const need = std::to_string(n) + " integrals at max_regions " +
                           std::to_string(cap);
  if (off > kScratchBytes)
    return fail(Y3G_ERR_ARG, "y3g_integrate_noalloc: " + need + " need more than its " +
                               std::to_string(kScratchBytes) + " bytes of scratch");
  if (!slab) return 0;
  // per workgroup: lo, hi [6][cap], est, err, raw [cap] (doubles), axis [cap]
  size_t const per_slot = (size_t)cap * (2 * kMaxDim * sizeof(double) + 3 * sizeof(double) + 1);
  size_t const fixed = 6 * 256 + 64;
  size_t const room = kScratchBytes - off;
  if (room <= fixed || (room - fixed) / per_slot < 1)
    return fail(Y3G_ERR_ARG, "y3g_integrate_noalloc: " + need + " need more than its " +
                               std::to_string(kScratchBytes) + " bytes of scratch (one "
                               "workgroup's region slab is " + std::to_string(per_slot) +
                               " bytes)");
  size_t const slots = std::min<size_t>((room - fixed) / per_slot, 1 << 20);
  size_t const nr = slots * (size_t)cap;
  B.lo = reinterpret_cast<double*>(take(nr * kMaxDim * sizeof(double)));
  B.hi = reinterpret_cast<double*>(take(nr * kMaxDim * sizeof(double)));
  B.est = reinterpret_cast<double*>(take(nr * sizeof(double)));
  B.err = reinterpret_cast<double*>(take(nr * sizeof(double)));
  B.raw = reinterpret_cast<double*>(take(nr * sizeof(double)));
  B.ax = reinterpret_cast<int8_t*>(take(nr + 64));
  if (off > kScratchBytes)  // cannot happen: slots were sized to fit
    return fail(Y3G_ERR_STATE, "y3g_integrate_noalloc: scratch layout");
  B.slab_slots = (int)slots;
  return 0;
}

static int run_group(y3g_runtime* rt, const y3g_integral* igs,
                     const std::vector<int>& ids, const y3g_options& o,
                     int algorithm, KernArgs A, const CallBufs& B)
{
  if (ids.empty()) return 0;
  std::vector<int> order;
  lpt_order(igs, ids, order);
  HIPCHK(hipMemcpyAsync(B.order, order.data(), sizeof(int) * order.size(),
                        hipMemcpyHostToDevice, rt->stream));
  if (int rc = zero_words(rt, B, B.queue, 1)) return rc;
  A.n = (int)order.size();
  A.order = B.order;
  A.queue = B.queue;
  int const block = o.block_threads == 256 ? 256 : 512;
  int slots = o.single ? 1 : resident_slots(rt, o, algorithm);
  slots = std::max(1, std::min<int>(slots, (int)ids.size()));
  if (algorithm == Y3G_PAGANI) {
    if (B.noalloc) {
      slots = std::min(slots, B.slab_slots);
      A.slab_lo = B.lo, A.slab_hi = B.hi;
      A.slab_est = B.est, A.slab_err = B.err, A.slab_raw = B.raw;
      A.slab_ax = B.ax;
    } else {
      if (int rc = ensure_slab(rt, std::max(slots, rt->slab_slots), A.cap))
        return rc;
      A.slab_lo = rt->slab_lo, A.slab_hi = rt->slab_hi;
      A.slab_est = rt->slab_est, A.slab_err = rt->slab_err, A.slab_raw = rt->slab_raw;
      A.slab_ax = rt->slab_ax;
    }
    HIPCHK(launch_cubature(A, o.precision, block, slots, rt->stream));
  } else {
    HIPCHK(launch_vegas(A, o.precision, block, slots, rt->stream));
  }
  return 0;
}

static int integrate_call(y3g_runtime* rt, const y3g_integral* igs, size_t n,
                          const y3g_options* opt, y3g_result* out, bool noalloc)
{
  char const* const who = noalloc ? "y3g_integrate_noalloc" : "y3g_integrate";
  if (!rt || (!igs && n) || (!out && n))
    return fail(Y3G_ERR_ARG, std::string(who) + ": bad arguments");
  y3g_options o;
  y3g_options_default(&o);
  if (opt) o = *opt;
  if (n == 0) return 0;
  if (rt->n_points <= 0)
    return fail(Y3G_ERR_STATE, std::string(who) + ": no tables (call y3g_tables_build)");
  if (o.max_regions < 64 || o.max_split_per_step < 1 ||
      o.max_split_per_step > 1024 || o.vegas_nbins < 1 || o.vegas_nbins > 128 ||
      o.vegas_calls < 2 || o.vegas_itmax < 1)
    return fail(Y3G_ERR_ARG, std::string(who) + ": options out of range");
  if (o.n_bands < 0 || o.n_bands > 8)
    return fail(Y3G_ERR_ARG, std::string(who) + ": n_bands out of range");
  for (int k = 0; k < o.n_bands; ++k)
    if (o.band_regions[k] < 64 || o.band_regions[k] > o.max_regions ||
        (k > 0 && !(o.band_eps[k] > o.band_eps[k - 1])))
      return fail(Y3G_ERR_ARG, std::string(who) +
                                 ": tolerance bands need ascending edges and "
                                 "budgets in [64, max_regions]");
  if (o.ext_regions < 0 || o.ext_regions > o.max_regions || !(o.ext_ratio >= 0.0))
    return fail(Y3G_ERR_ARG, std::string(who) +
                               ": ext_regions must lie in [0, max_regions] and ext_ratio >= 0");
  if (noalloc && o.single)
    return fail(Y3G_ERR_ARG, "y3g_integrate_noalloc: the plain path (single) allocates its "
                             "own workspace; use y3g_integrate");
  // tolerance-band budgets become per-integral call limits
  std::vector<y3g_integral> banded;
  if (o.n_bands > 0) {
    banded.assign(igs, igs + n);
    for (auto& g : banded) {
      if (g.algorithm == Y3G_MCUBES) continue;
      for (int k = 0; k < o.n_bands; ++k)
        if (g.eps_rel <= o.band_eps[k]) {
          long long const lim = 2LL * o.band_regions[k] * gm_num_nodes(g.ndim);
          if (lim < g.max_eval) g.max_eval = lim;
          break;
        }
    }
    igs = banded.data();
  }
  std::vector<int> cub, veg;
  bool any_des = false;
  for (size_t i = 0; i < n; ++i) {
    y3g_integral const& g = igs[i];
    if (g.ndim < 2 || g.ndim > kMaxDim || g.point < 0 ||
        g.point >= rt->n_points || g.kind < 0 || g.kind > Y3G_POP_J ||
        !(g.eps_rel >= 0) || g.max_eval < 1)
      return fail(Y3G_ERR_ARG, std::string(who) + ": integral " + std::to_string(i) +
                                 " is malformed");
    if (is_des_kind(g.kind)) {
      if (noalloc)
        return fail(Y3G_ERR_ARG, "y3g_integrate_noalloc: integral " + std::to_string(i) +
                                   ": the real-data kinds upload their DES data on demand; "
                                   "use y3g_integrate");
      if (g.ndim != des_ndim(g.kind) || o.precision != Y3G_FP64 || o.single)
        return fail(Y3G_ERR_ARG, "y3g_integrate: integral " + std::to_string(i) +
                                   ": the real-data integrands run on the FP64 batched "
                                   "integrators only (precision fp64)");
      if (!has_des(rt, g.point))
        return fail(Y3G_ERR_STATE, "y3g_integrate: integral " + std::to_string(i) +
                                     ": its point has no DES data (y3g_des_upload)");
      std::vector<double> const& dd = rt->des[g.point];
      long long const nt = (long long)dd.size() - kDesSig;
      if ((g.kind == Y3G_DSIGMA6D_DES || g.kind == Y3G_DSIGMA3D_DES) &&
          !(nt > 0 && sigmax_ok(dd.data() + kDesSig, nt)))
        return fail(Y3G_ERR_STATE, "y3g_integrate: integral " + std::to_string(i) +
                                     ": its point's DES data has no SIG_MAX tables");
      if (g.kind == Y3G_COUNT4D_CARD && !(nt > 0 && lclt_ok(dd.data() + kDesSig, nt)))
        return fail(Y3G_ERR_STATE, "y3g_integrate: integral " + std::to_string(i) +
                                     ": its point's DES data has no INT_LC_LT_DES tables");
      if (is_pop_kind(g.kind) && !(nt > 0 && pop_ok(dd.data() + kDesSig, nt)))
        return fail(Y3G_ERR_STATE, "y3g_integrate: integral " + std::to_string(i) +
                                     ": its point's data has no P[X] operator tables");
      any_des = true;
    }
    for (int a = 0; a < g.ndim; ++a)
      if (!(g.highs[a] > g.lows[a]))
        return fail(Y3G_ERR_ARG, std::string(who) + ": empty volume in integral " +
                                   std::to_string(i));
    (g.algorithm == Y3G_MCUBES ? veg : cub).push_back((int)i);
  }
  HIPCHK(hipSetDevice(rt->device));
  CallBufs B;
  if (noalloc) {
    if (int rc = carve_scratch(rt, n, o.max_regions, !cub.empty(), B)) return rc;
  } else {
    if (any_des)
      if (int rc = ensure_des(rt)) return rc;
    if (int rc = ensure_batch(rt, n)) return rc;
    B.igs = rt->d_igs, B.res = rt->d_res, B.order = rt->d_order;
    B.queue = rt->d_queue, B.nevals = rt->d_nevals;
  }
  HIPCHK(hipMemcpyAsync(B.igs, igs, sizeof(y3g_integral) * n,
                        hipMemcpyHostToDevice, rt->stream));
  if (int rc = zero_words(rt, B, B.nevals, 2)) return rc;
  KernArgs A;
  std::memset(&A, 0, sizeof(A));
  A.igs = B.igs;
  A.points = rt->points;
  A.hmf32 = rt->hmf32, A.hmf64 = rt->hmf64, A.hmf_stride = rt->hmf_stride;
  A.da32 = rt->da32, A.da64 = rt->da64, A.da_stride = rt->da_stride;
  A.sci32 = rt->sci32, A.sci64 = rt->sci64, A.sci_stride = rt->sci_stride;
  if (o.hmf_smooth && smooth_ok(rt->spec))
    A.hmfc32 = rt->hmfc32, A.hmfc64 = rt->hmfc64, A.hmfc_stride = rt->hmfc_stride;
  A.des = any_des ? rt->d_des : nullptr;
  A.des_off = any_des ? rt->d_des_off : nullptr;
  A.cap = o.max_regions;
  A.kmax = o.max_split_per_step;
  A.init_split = o.init_split;
  A.init_lnm = o.init_split_lnm;
  A.err_scale = o.err_scale;
  A.ext_ratio = o.ext_ratio;
  A.ext_regions = o.ext_regions;
  A.err_c = o.err_c;
  A.err_smin = o.err_smin;
  A.tl_c = o.tl_c;
  A.tl_smin = o.tl_smin;
  A.vegas_calls = o.vegas_calls;
  A.vegas_itmax = o.vegas_itmax;
  A.vegas_itskip = o.vegas_itskip;
  A.vegas_nbins = o.vegas_nbins;
  A.vegas_alpha = o.vegas_alpha;
  A.queue = B.queue;
  A.out = B.res;
  A.nevals = B.nevals;
  HIPCHK(hipEventRecord(rt->ev0, rt->stream));
  long long extra_nev = 0;
  if (o.single) {
    // plain path: one integral per launch sequence, serial over integrals;
    // cubature regions spread over the whole device
    if (!cub.empty())
      if (int rc = ensure_single(rt, o.max_regions)) return rc;
    std::vector<y3g_result> host(n);
    long long nev_single = 0;
    for (size_t i = 0; i < n; ++i) {
      if (igs[i].algorithm == Y3G_MCUBES) {
        std::vector<int> one{(int)i};
        if (int rc = run_group(rt, igs, one, o, Y3G_MCUBES, A, B)) return rc;
        continue;
      }
      SingleArgs sa = rt->sg;
      sa.K = A;
      sa.ig = rt->d_one;
      HIPCHK(hipMemcpyAsync(rt->d_one, &igs[i], sizeof(y3g_integral),
                            hipMemcpyHostToDevice, rt->stream));
      int const grid = rt->n_cu * (o.precision == Y3G_FP64 ? 1 : 2);
      HIPCHK(run_single(sa, igs[i], o.precision,
                        std::max(o.max_split_per_step, kSingleKmax), o.init_split,
                        o.init_split_lnm, grid,
                        rt->stream, &host[i], &nev_single));
    }
    // the VEGAS results were written on the device; cubature ones on the host
    HIPCHK(hipStreamSynchronize(rt->stream));
    std::vector<y3g_result> dres(n);
    HIPCHK(hipMemcpy(dres.data(), B.res, sizeof(y3g_result) * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i)
      if (igs[i].algorithm != Y3G_MCUBES) dres[i] = host[i];
    HIPCHK(hipMemcpy(B.res, dres.data(), sizeof(y3g_result) * n, hipMemcpyHostToDevice));
    extra_nev = nev_single;
  } else {
    if (int rc = run_group(rt, igs, cub, o, Y3G_PAGANI, A, B)) return rc;
    if (int rc = run_group(rt, igs, veg, o, Y3G_MCUBES, A, B)) return rc;
  }
  HIPCHK(hipEventRecord(rt->ev1, rt->stream));
  HIPCHK(hipMemcpyAsync(out, B.res, sizeof(y3g_result) * n,
                        hipMemcpyDeviceToHost, rt->stream));
  unsigned long long nev = 0;
  HIPCHK(hipMemcpyAsync(&nev, B.nevals, sizeof(nev), hipMemcpyDeviceToHost,
                        rt->stream));
  HIPCHK(hipStreamSynchronize(rt->stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, rt->ev0, rt->ev1));
  rt->last_ms = ms;
  rt->last_nevals = (long long)nev + extra_nev;
  return 0;
}

int y3g_integrate(y3g_runtime* rt, const y3g_integral* igs, size_t n,
                  const y3g_options* opt, y3g_result* out)
{
  return integrate_call(rt, igs, n, opt, out, false);
}

int y3g_integrate_noalloc(y3g_runtime* rt, const y3g_integral* igs, size_t n,
                          const y3g_options* opt, y3g_result* out)
{
  return integrate_call(rt, igs, n, opt, out, true);
}

int y3g_runtime_footprint(y3g_runtime* rt, const y3g_options* opt, size_t* bytes)
{
  if (!rt || !bytes) return fail(Y3G_ERR_ARG, "y3g_runtime_footprint: bad arguments");
  y3g_options o;
  y3g_options_default(&o);
  if (opt) o = *opt;
  size_t const per_region = 2 * kMaxDim * sizeof(double) + 3 * sizeof(double) + 1;
  size_t const slots = (size_t)resident_slots(rt, o, Y3G_PAGANI);
  *bytes = slots * (size_t)std::max(o.max_regions, 1) * per_region;
  return 0;
}

int y3g_runtime_memory(y3g_runtime* rt, size_t* free_bytes, size_t* total_bytes,
                       size_t* runtime_bytes)
{
  if (!rt) return fail(Y3G_ERR_ARG, "y3g_runtime_memory: null runtime");
  HIPCHK(hipSetDevice(rt->device));
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  if (free_bytes) *free_bytes = fr;
  if (total_bytes) *total_bytes = tot;
  if (runtime_bytes) {
    size_t const np = (size_t)std::max(rt->n_points, 0);
    size_t b = np * (sizeof(PointDev) + (size_t)rt->hmf_stride * 12 +
                     (size_t)rt->hmfc_stride * 12 + (size_t)rt->da_stride * 12 +
                     (size_t)rt->sci_stride * 12);
    b += (size_t)rt->slab_slots * rt->slab_cap * (2 * kMaxDim * 8 + 3 * 8 + 1);
    b += rt->batch_cap * (sizeof(y3g_integral) + sizeof(y3g_result) + sizeof(int));
    b += (size_t)rt->sg_cap * (2 * kMaxDim * 8 + 6 * 8);
    b += (rt->des_cap + rt->seldata_cap + rt->selnodes_cap + rt->selout_cap) * 8;
    b += (size_t)rt->nfw_nx * rt->nfw_nm * 8;
    *runtime_bytes = b;
  }
  return 0;
}

int y3g_last_timing(y3g_runtime* rt, double* kernel_ms, long long* nevals)
{
  if (!rt) return fail(Y3G_ERR_ARG, "null runtime");
  if (kernel_ms) *kernel_ms = rt->last_ms;
  if (nevals) *nevals = rt->last_nevals;
  return 0;
}

int y3g_eval_functor(y3g_runtime* rt, int which, int precision,
                     const double* in, int nin, size_t n, const double* par,
                     int npar, const y3g_integral* aux, double* out)
{
  if (!rt || !in || !out || nin < 1 || which < 0 || which > Y3G_F_SCI_RZ)
    return fail(Y3G_ERR_ARG, "y3g_eval_functor: bad arguments");
  if (which == Y3G_F_SCI_AVERAGE) {
    // par = {n, z[n], sci[n]}, z increasing
    double const a = par && npar >= 1 ? par[0] : 0.0;
    if (nin != 1 || !(a >= 2.0 && a <= 65536.0) || a != double(int(a)) ||
        (long long)npar != 1 + 2LL * (long long)a)
      return fail(Y3G_ERR_ARG, "y3g_eval_functor: SCI_AVERAGE takes (zl) and "
                               "{n, zlense[n], sci_average[n]}");
    for (int k = 1; k < int(a); ++k)
      if (!(par[1 + k] > par[k]))
        return fail(Y3G_ERR_ARG, "y3g_eval_functor: SCI_AVERAGE z knots not increasing");
  }
  if (which == Y3G_F_SCI_RZ) {
    // par = {n_r, n_z, r[n_r], z[n_z], table[n_z][n_r]}
    double const a = par && npar >= 2 ? par[0] : 0.0, b = par && npar >= 2 ? par[1] : 0.0;
    if (nin != 2 || !(a >= 2.0 && b >= 2.0 && a <= 65536.0 && b <= 65536.0) ||
        a != double(int(a)) || b != double(int(b)) ||
        (long long)npar != 2 + (long long)(a + b) + (long long)a * (long long)b)
      return fail(Y3G_ERR_ARG, "y3g_eval_functor: SCI_RZ takes (r, zt) and "
                               "{n_r, n_z, r, z, sigma_crit_inv[n_z][n_r]}");
    for (int k = 1; k < int(a); ++k)
      if (!(par[2 + k] > par[1 + k]))
        return fail(Y3G_ERR_ARG, "y3g_eval_functor: SCI_RZ r knots not increasing");
    for (int k = 1; k < int(b); ++k)
      if (!(par[2 + int(a) + k] > par[1 + int(a) + k]))
        return fail(Y3G_ERR_ARG, "y3g_eval_functor: SCI_RZ z knots not increasing");
  }
  if (which == Y3G_F_LC_LT && nin != 3)
    return fail(Y3G_ERR_ARG, "y3g_eval_functor: LC_LT takes (lc, lt, zt)");
  if (which == Y3G_F_NFW_DSIGMA_MIS) {
    if (nin != 3 || !par || npar != 2 || precision != Y3G_FP64)
      return fail(Y3G_ERR_ARG, "y3g_eval_functor: NFW_DSIGMA_MIS takes (r, rmis, lnM), "
                               "{c, rho} and FP64");
    if (!rt->d_nfwmis)
      return fail(Y3G_ERR_STATE, "y3g_eval_functor: NFW_DSIGMA_MIS needs its table "
                                 "(y3g_nfw_mis_table_build / _upload)");
  }
  if (which == Y3G_F_MOR_DES_LOG && (nin != 3 || !par || npar != 6))
    return fail(Y3G_ERR_ARG, "y3g_eval_functor: MOR_DES_LOG takes (lt, lnM, zt) and "
                             "{log10 Mmin, log10 ratio, alpha, sigma_intr, epsilon, z_pivot}");
  if (which == Y3G_F_SIG_MAX) {
    // par = {n_r, n_m, n_z, r, lnM, z, sigma_1, sigma_2, bias}
    if (nin != 3 || !par)
      return fail(Y3G_ERR_ARG, "y3g_eval_functor: SIG_MAX takes (r, lnM, zt) and tables");
    if (!sigmax_ok(par, npar))
      return fail(Y3G_ERR_ARG, "y3g_eval_functor: SIG_MAX table sizes or axes");
  }
  if (which == Y3G_F_EMG_DES_PDF || which == Y3G_F_EMG_DES_CDF) {
    // par = {n, z[n], eight parameter curves of n values}
    if (nin != 3 || !par || npar < 1)
      return fail(Y3G_ERR_ARG, "y3g_eval_functor: EMG_DES takes (lob, ltr, z) and curves");
    double const a = par[0];
    if (!(a >= 2.0 && a <= 65536.0) || a != double(int(a)) ||
        (long long)npar != 1 + 9LL * (long long)a)
      return fail(Y3G_ERR_ARG, "y3g_eval_functor: EMG_DES curve sizes");
    for (int k = 1; k < int(a); ++k)
      if (!(par[1 + k] > par[k]))
        return fail(Y3G_ERR_ARG, "y3g_eval_functor: EMG_DES z knots not increasing");
  }
  if (which == Y3G_F_MOR_DES && (nin != 3 || !par || npar != 6))
    return fail(Y3G_ERR_ARG, "y3g_eval_functor: MOR_DES takes (lt, lnM, zt) and "
                             "{A, B, alpha, sigma_intr, epsilon, z_pivot}");
  if (which == Y3G_F_INT_LC_LT_DES) {
    // par = {n_lt, n_zt, lt knots, zt knots, four n_lt x n_zt tables}
    if (nin != 3 || !par || npar < 2)
      return fail(Y3G_ERR_ARG, "y3g_eval_functor: INT_LC_LT_DES takes (lc, lt, zt) and tables");
    double const a = par[0], b = par[1];
    if (!(a >= 2.0 && b >= 2.0 && a <= 4096.0 && b <= 4096.0) || a != double(int(a)) ||
        b != double(int(b)) || (long long)npar != 2 + (long long)(a + b) + 4LL * (long long)(a * b))
      return fail(Y3G_ERR_ARG, "y3g_eval_functor: INT_LC_LT_DES table sizes");
    for (int k = 1; k < int(a); ++k)
      if (!(par[2 + k] > par[1 + k]))
        return fail(Y3G_ERR_ARG, "y3g_eval_functor: INT_LC_LT_DES lt knots not increasing");
    for (int k = 1; k < int(b); ++k)
      if (!(par[2 + int(a) + k] > par[1 + int(a) + k]))
        return fail(Y3G_ERR_ARG, "y3g_eval_functor: INT_LC_LT_DES zt knots not increasing");
  }
  if (n == 0) return 0;
  HIPCHK(hipSetDevice(rt->device));
  bool const needs_tables = which == Y3G_F_DV_DO_DZ || which == Y3G_F_HMF ||
                            which == Y3G_F_MOR || which == Y3G_F_LO_LT_PDF ||
                            which == Y3G_F_LO_LT_BIN || which == Y3G_F_SCI ||
                            which == Y3G_F_INTEGRAND || which == Y3G_F_INTEGRAND_PAIRED ||
                            which == Y3G_F_INTEGRAND_NODE;
  if (needs_tables && rt->n_points <= 0)
    return fail(Y3G_ERR_STATE, "y3g_eval_functor: no tables");
  bool const is_ig = which == Y3G_F_INTEGRAND || which == Y3G_F_INTEGRAND_PAIRED ||
                     which == Y3G_F_INTEGRAND_NODE;
  if (is_ig && (!aux || aux->ndim != nin))
    return fail(Y3G_ERR_ARG, "y3g_eval_functor: integrand needs aux integral");
  int const pt = is_ig ? aux->point : 0;
  if (needs_tables && (pt < 0 || pt >= rt->n_points))
    return fail(Y3G_ERR_ARG, "y3g_eval_functor: bad point");
  bool const des_ig = is_ig && is_des_kind(aux->kind);
  bool des_tab_ok = des_ig && has_des(rt, pt);
  if (des_tab_ok) {
    std::vector<double> const& dd = rt->des[pt];
    long long const nt = (long long)dd.size() - kDesSig;
    if (aux->kind == Y3G_DSIGMA6D_DES || aux->kind == Y3G_DSIGMA3D_DES)
      des_tab_ok = nt > 0 && sigmax_ok(dd.data() + kDesSig, nt);
    else if (aux->kind == Y3G_COUNT4D_CARD)
      des_tab_ok = nt > 0 && lclt_ok(dd.data() + kDesSig, nt);
    else if (is_pop_kind(aux->kind))
      des_tab_ok = nt > 0 && pop_ok(dd.data() + kDesSig, nt);
  }
  if (des_ig && (which != Y3G_F_INTEGRAND || precision != Y3G_FP64 || !des_tab_ok))
    return fail(Y3G_ERR_ARG, "y3g_eval_functor: the real-data Delta-Sigma integrand is "
                             "Y3G_F_INTEGRAND in FP64 with the point's DES data");
  if (des_ig)
    if (int rc = ensure_des(rt)) return rc;
  long long des_off = 0;
  if (des_ig) HIPCHK(hipMemcpy(&des_off, rt->d_des_off + pt, sizeof(des_off),
                               hipMemcpyDeviceToHost));
  double *d_in = nullptr, *d_out = nullptr, *d_par = nullptr;
  y3g_integral* d_ig = nullptr;
  int rc = 0;
  if ((rc = dalloc(&d_in, n * nin)) || (rc = dalloc(&d_out, n)) ||
      (rc = dalloc(&d_par, npar > 0 ? npar : 1)) || (rc = dalloc(&d_ig, 1))) {
    dfree(d_in), dfree(d_out), dfree(d_par), dfree(d_ig);
    return rc;
  }
  auto cleanup = [&]() { dfree(d_in), dfree(d_out), dfree(d_par), dfree(d_ig); };
  if (hipMemcpy(d_in, in, sizeof(double) * n * nin, hipMemcpyHostToDevice) ||
      (npar > 0 &&
       hipMemcpy(d_par, par, sizeof(double) * npar, hipMemcpyHostToDevice)) ||
      (aux && hipMemcpy(d_ig, aux, sizeof(*aux), hipMemcpyHostToDevice))) {
    cleanup();
    return fail(Y3G_ERR_HIP, "y3g_eval_functor: copy in");
  }
  FunctorArgs F;
  F.which = which;
  F.in = d_in;
  F.nin = nin;
  F.n = n;
  F.par = d_par;
  F.points = rt->points ? rt->points + (needs_tables ? pt : 0) : nullptr;
  bool const f64 = precision == Y3G_FP64;
  if (rt->n_points > 0) {
    F.hmf = f64 ? (const void*)(rt->hmf64 + (size_t)pt * rt->hmf_stride)
                : (const void*)(rt->hmf32 + (size_t)pt * rt->hmf_stride);
    F.da = f64 ? (const void*)(rt->da64 + (size_t)pt * rt->da_stride)
               : (const void*)(rt->da32 + (size_t)pt * rt->da_stride);
    F.sci = f64 ? (const void*)(rt->sci64 + (size_t)pt * rt->sci_stride)
                : (const void*)(rt->sci32 + (size_t)pt * rt->sci_stride);
  } else {
    F.hmf = F.da = F.sci = nullptr;
  }
  F.des = des_ig ? rt->d_des + des_off : nullptr;
  F.ig = d_ig;
  F.nfwmis = rt->d_nfwmis;
  F.nfw_nx = rt->nfw_nx;
  F.nfw_nm = rt->nfw_nm;
  F.out = d_out;
  hipError_t e = launch_functor(F, precision, rt->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(rt->stream);
  if (e == hipSuccess)
    e = hipMemcpy(out, d_out, sizeof(double) * n, hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess)
    return fail(Y3G_ERR_HIP, std::string("y3g_eval_functor: ") + hipGetErrorString(e));
  return 0;
}

int y3g_selfn_upload(y3g_runtime* rt, int point, const double* lnm, int n_lnm,
                     const double* z, int n_z, const double* s_stack, int n_bins)
{
  if (!rt || point < 0 || !lnm || !z || !s_stack || n_lnm < 2 || n_z < 2 || n_bins < 1)
    return fail(Y3G_ERR_ARG, "y3g_selfn_upload: bad arguments");
  for (int i = 1; i < n_lnm; ++i)
    if (!(lnm[i] > lnm[i - 1])) return fail(Y3G_ERR_ARG, "y3g_selfn_upload: lnM axis not increasing");
  for (int i = 1; i < n_z; ++i)
    if (!(z[i] > z[i - 1])) return fail(Y3G_ERR_ARG, "y3g_selfn_upload: z axis not increasing");
  if ((size_t)point >= rt->sel.size()) rt->sel.resize((size_t)point + 1);
  auto& h = rt->sel[point];
  h.lnm.assign(lnm, lnm + n_lnm);
  h.z.assign(z, z + n_z);
  h.s.assign(s_stack, s_stack + (size_t)n_bins * n_z * n_lnm);
  h.n_bins = n_bins;
  rt->sel_dirty = true;
  return 0;
}

int y3g_des_upload(y3g_runtime* rt, int point, const double* des, int n_des)
{
  if (!rt || point < 0 || !des || n_des < kDesSig)
    return fail(Y3G_ERR_ARG, "y3g_des_upload: bad arguments");
  for (int k = 0; k < kDesSig; ++k)
    if (!std::isfinite(des[k]))
      return fail(Y3G_ERR_ARG, "y3g_des_upload: non-finite parameter " + std::to_string(k));
  if (n_des > kDesSig && !sigmax_ok(des + kDesSig, n_des - kDesSig) &&
      !lclt_ok(des + kDesSig, n_des - kDesSig) && !pop_ok(des + kDesSig, n_des - kDesSig))
    return fail(Y3G_ERR_ARG, "y3g_des_upload: SIG_MAX, INT_LC_LT_DES or P[X] table sizes "
                             "or axes");
  if ((size_t)point >= rt->des.size()) rt->des.resize((size_t)point + 1);
  rt->des[point].assign(des, des + n_des);
  rt->des_dirty = true;
  return 0;
}

int y3g_selgl_counts(y3g_runtime* rt, const y3g_selgl_spec* spec, int n_points,
                     int max_bins, double* out)
{
  if (!rt || !spec || !out || n_points < 1 || max_bins < 1)
    return fail(Y3G_ERR_ARG, "y3g_selgl_counts: bad arguments");
  if (spec->n_lnm < 1 || spec->n_lnm > Y3G_SELGL_MAX_NODES || spec->n_z < 1 ||
      spec->n_z > Y3G_SELGL_MAX_NODES || !(spec->lnm_hi > spec->lnm_lo) ||
      !(spec->zt_hi > spec->zt_lo))
    return fail(Y3G_ERR_ARG, "y3g_selgl_counts: bad node counts or windows");
  if (n_points > rt->n_points)
    return fail(Y3G_ERR_STATE, "y3g_selgl_counts: points without tables");
  if ((size_t)n_points > rt->sel.size())
    return fail(Y3G_ERR_STATE, "y3g_selgl_counts: points without a selection function");
  for (int p = 0; p < n_points; ++p)
    if (rt->sel[p].n_bins < 1)
      return fail(Y3G_ERR_STATE, "y3g_selgl_counts: points without a selection function");
  HIPCHK(hipSetDevice(rt->device));
  if (rt->sel_dirty) {
    std::vector<double> data;
    std::vector<SelPoint> pts(rt->sel.size());
    for (size_t p = 0; p < rt->sel.size(); ++p) {
      auto const& h = rt->sel[p];
      SelPoint& sp = pts[p];
      sp.n_lnm = int(h.lnm.size());
      sp.n_z = int(h.z.size());
      sp.n_bins = h.n_bins;
      sp.off_lnm = (long long)data.size();
      data.insert(data.end(), h.lnm.begin(), h.lnm.end());
      sp.off_z = (long long)data.size();
      data.insert(data.end(), h.z.begin(), h.z.end());
      sp.off_s = (long long)data.size();
      data.insert(data.end(), h.s.begin(), h.s.end());
    }
    if (data.size() > rt->seldata_cap) {  // grow only: a sampler re-uploads every sample
      dfree(rt->d_seldata);
      rt->seldata_cap = 0;
      if (int rc = dalloc(&rt->d_seldata, data.size())) return rc;
      rt->seldata_cap = data.size();
    }
    if (pts.size() > rt->selpts_cap) {
      dfree(rt->d_selpts);
      rt->selpts_cap = 0;
      if (int rc = dalloc(&rt->d_selpts, pts.size())) return rc;
      rt->selpts_cap = pts.size();
    }
    HIPCHK(hipMemcpy(rt->d_seldata, data.data(), sizeof(double) * data.size(),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(rt->d_selpts, pts.data(), sizeof(SelPoint) * pts.size(),
                     hipMemcpyHostToDevice));
    rt->sel_dirty = false;
  }
  // GL nodes and weights: computed and uploaded when the spec changes
  if (!rt->selnodes_valid || std::memcmp(&rt->selnodes_spec, spec, sizeof(*spec)) != 0) {
    std::vector<double> nodes(2 * (size_t)(spec->n_lnm + spec->n_z));
    gl_nodes(spec->n_lnm, spec->lnm_lo, spec->lnm_hi, nodes.data(),
             nodes.data() + spec->n_lnm);
    gl_nodes(spec->n_z, spec->zt_lo, spec->zt_hi, nodes.data() + 2 * spec->n_lnm,
             nodes.data() + 2 * spec->n_lnm + spec->n_z);
    rt->selnodes_valid = false;
    if (nodes.size() > rt->selnodes_cap) {
      dfree(rt->d_selnodes);
      rt->selnodes_cap = 0;
      if (int rc = dalloc(&rt->d_selnodes, nodes.size())) return rc;
      rt->selnodes_cap = nodes.size();
    }
    HIPCHK(hipMemcpy(rt->d_selnodes, nodes.data(), sizeof(double) * nodes.size(),
                     hipMemcpyHostToDevice));
    rt->selnodes_spec = *spec;
    rt->selnodes_valid = true;
  }
  size_t const nout = (size_t)n_points * max_bins;
  if (nout > rt->selout_cap) {
    dfree(rt->d_selout);
    if (int rc = dalloc(&rt->d_selout, nout)) return rc;
    rt->selout_cap = nout;
  }
  SelGLArgs A{};
  A.points = rt->points;
  A.hmf64 = rt->hmf64;
  A.hmf_stride = rt->hmf_stride;
  A.da64 = rt->da64;
  A.da_stride = rt->da_stride;
  A.sel = rt->d_selpts;
  A.data = rt->d_seldata;
  A.nodes = rt->d_selnodes;
  A.n_lnm = spec->n_lnm;
  A.n_z = spec->n_z;
  A.max_bins = max_bins;
  A.max_axes = 0;
  for (int p = 0; p < n_points; ++p) {
    size_t const nax = rt->sel[p].lnm.size() + rt->sel[p].z.size();
    A.max_axes = (int)std::min<size_t>(std::max<size_t>(A.max_axes, nax), 1u << 30);
  }
  A.out = rt->d_selout;
  HIPCHK(hipEventRecord(rt->ev0, rt->stream));
  HIPCHK(launch_selgl(A, n_points, rt->stream));
  HIPCHK(hipEventRecord(rt->ev1, rt->stream));
  HIPCHK(hipMemcpyAsync(out, rt->d_selout, sizeof(double) * nout, hipMemcpyDeviceToHost,
                        rt->stream));
  HIPCHK(hipStreamSynchronize(rt->stream));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, rt->ev0, rt->ev1));
  rt->last_ms = ms;
  rt->last_nevals = (long long)nout * spec->n_lnm * spec->n_z;
  return 0;
}

}  // extern "C"

// ------------------------------------------ NFW_DSIGMA_MIS table
static int nfw_mis_alloc(y3g_runtime* rt, int nx, int nm)
{
  if (rt->d_nfwmis && rt->nfw_nx == nx && rt->nfw_nm == nm) return 0;
  dfree(rt->d_nfwmis);
  rt->nfw_nx = rt->nfw_nm = 0;
  if (int rc = dalloc(&rt->d_nfwmis, (size_t)nx + nm + (size_t)nx * nm)) return rc;
  rt->nfw_nx = nx;
  rt->nfw_nm = nm;
  return 0;
}

int y3g_nfw_mis_table_build(y3g_runtime* rt)
{
  if (!rt) return fail(Y3G_ERR_ARG, "y3g_nfw_mis_table_build: null runtime");
  HIPCHK(hipSetDevice(rt->device));
  int const nx = kNfwMisNx, nm = kNfwMisNm;
  // the reference's axes: ln x over [ln 1e-3, ln 5e3], ln xmis over
  // [ln 1e-2, ln 5e2], uniform
  std::vector<double> ax((size_t)nx + nm);
  double const a = std::log(1e-3), b = std::log(5e3), c = std::log(1e-2), d = std::log(5e2);
  for (int i = 0; i < nx; ++i) ax[i] = a + i * ((b - a) / (nx - 1));
  for (int j = 0; j < nm; ++j) ax[nx + j] = c + j * ((d - c) / (nm - 1));
  if (int rc = nfw_mis_alloc(rt, nx, nm)) return rc;
  HIPCHK(hipMemcpyAsync(rt->d_nfwmis, ax.data(), sizeof(double) * ax.size(),
                        hipMemcpyHostToDevice, rt->stream));
  HIPCHK(launch_nfw_mis_table(rt->d_nfwmis, nx, nm, rt->stream));
  HIPCHK(hipStreamSynchronize(rt->stream));
  return 0;
}

int y3g_nfw_mis_table_upload(y3g_runtime* rt, const double* logx, int nx,
                             const double* logxmis, int nm, const double* log_dsigma)
{
  if (!rt || !logx || !logxmis || !log_dsigma || nx < 2 || nm < 2 || nx > (1 << 16) ||
      nm > (1 << 16))
    return fail(Y3G_ERR_ARG, "y3g_nfw_mis_table_upload: bad arguments");
  for (int i = 1; i < nx; ++i)
    if (!(logx[i] > logx[i - 1]))
      return fail(Y3G_ERR_ARG, "y3g_nfw_mis_table_upload: ln x not increasing");
  for (int j = 1; j < nm; ++j)
    if (!(logxmis[j] > logxmis[j - 1]))
      return fail(Y3G_ERR_ARG, "y3g_nfw_mis_table_upload: ln xmis not increasing");
  for (size_t k = 0; k < (size_t)nx * nm; ++k)
    if (!std::isfinite(log_dsigma[k]))
      return fail(Y3G_ERR_ARG, "y3g_nfw_mis_table_upload: table value not finite");
  HIPCHK(hipSetDevice(rt->device));
  if (int rc = nfw_mis_alloc(rt, nx, nm)) return rc;
  HIPCHK(hipMemcpy(rt->d_nfwmis, logx, sizeof(double) * nx, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(rt->d_nfwmis + nx, logxmis, sizeof(double) * nm, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(rt->d_nfwmis + nx + nm, log_dsigma, sizeof(double) * nx * nm,
                   hipMemcpyHostToDevice));
  return 0;
}

int y3g_nfw_mis_table_shape(y3g_runtime* rt, int* nx, int* nm)
{
  if (!rt || !nx || !nm) return fail(Y3G_ERR_ARG, "y3g_nfw_mis_table_shape: bad arguments");
  *nx = rt->nfw_nx;
  *nm = rt->nfw_nm;
  return 0;
}

int y3g_nfw_mis_table_download(y3g_runtime* rt, double* logx, double* logxmis,
                               double* log_dsigma)
{
  if (!rt) return fail(Y3G_ERR_ARG, "y3g_nfw_mis_table_download: null runtime");
  if (!rt->d_nfwmis) return fail(Y3G_ERR_STATE, "y3g_nfw_mis_table_download: no table");
  HIPCHK(hipSetDevice(rt->device));
  int const nx = rt->nfw_nx, nm = rt->nfw_nm;
  if (logx) HIPCHK(hipMemcpy(logx, rt->d_nfwmis, sizeof(double) * nx, hipMemcpyDeviceToHost));
  if (logxmis)
    HIPCHK(hipMemcpy(logxmis, rt->d_nfwmis + nx, sizeof(double) * nm, hipMemcpyDeviceToHost));
  if (log_dsigma)
    HIPCHK(hipMemcpy(log_dsigma, rt->d_nfwmis + nx + nm, sizeof(double) * nx * nm,
                     hipMemcpyDeviceToHost));
  return 0;
}

// ------------------------------------------------------- streaming session
//
// The continuous-batching session of y3g_stream.hpp: ring, results and
// control block in coherent pinned host memory (the launch reads and writes
// them over the bus), the head counter, table generations and the slots'
// tables in device memory.

struct y3g_stream {
  y3g_runtime* rt = nullptr;
  y3g::StreamConfig cfg{};
  int workgroups = 0;
  size_t slab_bytes = 0, slot_bytes = 0;
  // pinned host memory
  StreamItem* ring = nullptr;
  y3g_result* res = nullptr;
  unsigned long long* seq = nullptr;
  StreamCtl* ctl = nullptr;
  PointDev* hpoints = nullptr;
  // device memory
  StreamDev* dev = nullptr;
  unsigned* ready = nullptr;
  // host state
  unsigned long long tail = 0;        // entries published
  unsigned long long low = 0;         // every entry below is finished and reported
  unsigned long long head_known = 0;  // entries claimed when the last launch ended
  std::vector<unsigned char> kind;    // per ring entry: 0 free, 1 integral, 2 table build
  std::vector<unsigned> gen;          // per slot: the last generation published
  std::vector<int> tables_pending;    // per slot: table builds not finished
  bool running = false;
  long long launches = 0;
  hipEvent_t done = nullptr;
  KernArgs A{};
  StreamArgs Z{};
};

namespace y3g {

static int stream_shape(y3g_runtime* rt, const StreamConfig& c, int* wg, size_t* slab,
                        size_t* slot)
{
  if (c.options.precision != Y3G_FP32SPLIT && c.options.precision != Y3G_FP64)
    return fail(Y3G_ERR_ARG, "stream: precision must be fp32split or fp64");
  if (c.options.max_regions < 64 || c.options.max_split_per_step < 1 ||
      c.options.max_split_per_step > 1024 || c.options.ext_regions < 0 ||
      c.options.ext_regions > c.options.max_regions || !(c.options.ext_ratio >= 0.0))
    return fail(Y3G_ERR_ARG, "stream: options out of range");
  if (int rc = check_spec(&c.spec)) return rc;
  HIPCHK(hipSetDevice(rt->device));
  int const per_cu = cubature_stream_blocks_per_cu(c.options.precision);
  if (per_cu < 1) return fail(Y3G_ERR_HIP, "stream: occupancy query failed");
  *wg = c.options.slots > 0 ? c.options.slots : rt->n_cu * per_cu;
  *slab = (size_t)*wg * c.options.max_regions * (2 * kMaxDim * 8 + 3 * 8 + 1);
  size_t const nh = (size_t)c.spec.n_lnm * c.spec.n_z;
  *slot = sizeof(PointDev) + nh * 12 + (size_t)c.spec.n_da * 12 + (size_t)c.spec.n_sci * 12 + 4;
  return 0;
}

int stream_footprint(y3g_runtime* rt, const StreamConfig& c, size_t* slab_bytes,
                     size_t* slot_bytes, int* workgroups)
{
  if (!rt || !slab_bytes || !slot_bytes || !workgroups)
    return fail(Y3G_ERR_ARG, "stream_footprint: bad arguments");
  return stream_shape(rt, c, workgroups, slab_bytes, slot_bytes);
}

static void stream_free(y3g_stream* s)
{
  if (s->ring) (void)hipHostFree(s->ring);
  if (s->res) (void)hipHostFree(s->res);
  if (s->seq) (void)hipHostFree(s->seq);
  if (s->ctl) (void)hipHostFree(s->ctl);
  if (s->hpoints) (void)hipHostFree(s->hpoints);
  dfree(s->dev);
  dfree(s->ready);
  if (s->done) (void)hipEventDestroy(s->done);
  delete s;
}

template <typename T> static hipError_t halloc(T** p, size_t n)
{
  // coherent: the launch polls and writes this memory while it runs
  hipError_t const e = hipHostMalloc(reinterpret_cast<void**>(p), sizeof(T) * n,
                                     hipHostMallocCoherent | hipHostMallocMapped);
  if (e == hipSuccess) std::memset(*p, 0, sizeof(T) * n);
  return e;
}

int stream_open(y3g_runtime* rt, const StreamConfig& c, y3g_stream** out)
{
  if (!rt || !out) return fail(Y3G_ERR_ARG, "stream_open: bad arguments");
  *out = nullptr;
  if (c.ring_items < 2 || c.table_slots < 1 || !(c.idle_ms > 0) || !(c.table_wait_ms > 0))
    return fail(Y3G_ERR_ARG, "stream_open: ring_items >= 2, table_slots >= 1 and positive "
                             "idle / wait bounds are needed");
  auto* s = new y3g_stream;
  s->rt = rt;
  s->cfg = c;
  auto bail = [&](int rc) {
    stream_free(s);
    return rc;
  };
  if (int rc = stream_shape(rt, c, &s->workgroups, &s->slab_bytes, &s->slot_bytes))
    return bail(rc);
  if (int rc = alloc_tables(rt, &c.spec, c.table_slots)) return bail(rc);
  if (int rc = ensure_slab(rt, s->workgroups, c.options.max_regions)) return bail(rc);
  size_t const n = (size_t)c.ring_items;
  if (halloc(&s->ring, n) != hipSuccess || halloc(&s->res, n) != hipSuccess ||
      halloc(&s->seq, n) != hipSuccess || halloc(&s->ctl, 1) != hipSuccess ||
      halloc(&s->hpoints, (size_t)c.table_slots) != hipSuccess)
    return bail(fail(Y3G_ERR_MEMORY, "stream_open: pinned host memory"));
  if (dalloc(&s->dev, 1) || dalloc(&s->ready, (size_t)c.table_slots))
    return bail(Y3G_ERR_MEMORY);
  if (hipMemset(s->dev, 0, sizeof(StreamDev)) != hipSuccess ||
      hipMemset(s->ready, 0, sizeof(unsigned) * c.table_slots) != hipSuccess ||
      hipEventCreateWithFlags(&s->done, hipEventDisableTiming) != hipSuccess)
    return bail(fail(Y3G_ERR_HIP, "stream_open: device set-up"));
  if (hipDeviceSynchronize() != hipSuccess)
    return bail(fail(Y3G_ERR_HIP, "stream_open: device set-up"));
  s->kind.assign(n, 0);
  s->gen.assign(c.table_slots, 0u);
  s->tables_pending.assign(c.table_slots, 0);
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, rt->device) != hipSuccess ||
      khz <= 0)
    return bail(fail(Y3G_ERR_HIP, "stream_open: wall-clock rate"));
  KernArgs& A = s->A;
  std::memset(&A, 0, sizeof(A));
  A.points = rt->points;
  A.hmf32 = rt->hmf32, A.hmf64 = rt->hmf64, A.hmf_stride = rt->hmf_stride;
  A.da32 = rt->da32, A.da64 = rt->da64, A.da_stride = rt->da_stride;
  A.sci32 = rt->sci32, A.sci64 = rt->sci64, A.sci_stride = rt->sci_stride;
  if (c.options.hmf_smooth && smooth_ok(c.spec))
    A.hmfc32 = rt->hmfc32, A.hmfc64 = rt->hmfc64, A.hmfc_stride = rt->hmfc_stride;
  A.slab_lo = rt->slab_lo, A.slab_hi = rt->slab_hi;
  A.slab_est = rt->slab_est, A.slab_err = rt->slab_err, A.slab_raw = rt->slab_raw;
  A.slab_ax = rt->slab_ax;
  A.cap = c.options.max_regions;
  A.kmax = c.options.max_split_per_step;
  A.init_split = c.options.init_split;
  A.init_lnm = c.options.init_split_lnm;
  A.err_scale = c.options.err_scale;
  A.ext_ratio = c.options.ext_ratio;
  A.ext_regions = c.options.ext_regions;
  A.err_c = c.options.err_c, A.err_smin = c.options.err_smin;
  A.tl_c = c.options.tl_c, A.tl_smin = c.options.tl_smin;
  A.nevals = rt->d_nevals;
  StreamArgs& Z = s->Z;
  Z.ring = s->ring;
  Z.res = s->res;
  Z.seq = s->seq;
  Z.ctl = s->ctl;
  Z.hpoints = s->hpoints;
  Z.dev = s->dev;
  Z.ready = s->ready;
  Z.points = rt->points;
  Z.tab.points = rt->points;
  Z.tab.spec = c.spec;
  Z.tab.hmf64 = rt->hmf64, Z.tab.hmf32 = rt->hmf32, Z.tab.hmf_stride = rt->hmf_stride;
  Z.tab.da64 = rt->da64, Z.tab.da32 = rt->da32, Z.tab.da_stride = rt->da_stride;
  Z.tab.sci64 = rt->sci64, Z.tab.sci32 = rt->sci32, Z.tab.sci_stride = rt->sci_stride;
  Z.tab.hmfc64 = rt->hmfc64, Z.tab.hmfc32 = rt->hmfc32, Z.tab.hmfc_stride = rt->hmfc_stride;
  Z.ring_cap = (unsigned)c.ring_items;
  Z.idle_ticks = (long long)(c.idle_ms * khz);
  Z.wait_ticks = (long long)(c.table_wait_ms * khz);
  if (hipMemset(rt->d_nevals, 0, 8) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return bail(fail(Y3G_ERR_HIP, "stream_open: device set-up"));
  *out = s;
  return 0;
}

size_t stream_room(const y3g_stream* s)
{
  return (size_t)s->cfg.ring_items - (size_t)(s->tail - s->low);
}

int stream_tables_pending(const y3g_stream* s, int slot)
{
  return slot >= 0 && slot < (int)s->tables_pending.size() ? s->tables_pending[slot] : 0;
}

// Make entries [tail, tail + n) visible to the launch.
static void stream_publish_tail(y3g_stream* s, unsigned long long n)
{
  s->tail += n;
  __atomic_store_n(&s->ctl->tail, s->tail, __ATOMIC_RELEASE);
}

int stream_tables(y3g_stream* s, int slot, const y3g_params& p, unsigned* gen)
{
  if (!s || !gen || slot < 0 || slot >= s->cfg.table_slots)
    return fail(Y3G_ERR_ARG, "stream_tables: bad arguments");
  if (stream_room(s) < 1) return fail(Y3G_ERR_STATE, "stream_tables: the ring is full");
  if (s->tables_pending[slot])
    return fail(Y3G_ERR_STATE, "stream_tables: the slot's last build has not finished");
  s->hpoints[slot] = make_point(p, s->cfg.spec);
  unsigned const g = ++s->gen[slot];
  size_t const k = (size_t)(s->tail % (unsigned long long)s->cfg.ring_items);
  StreamItem e;
  std::memset(&e, 0, sizeof(e));
  e.job = kStreamJobTables;
  e.slot = slot;
  e.gen = g;
  s->ring[k] = e;
  s->kind[k] = 2;
  s->tables_pending[slot] += 1;
  stream_publish_tail(s, 1);
  *gen = g;
  return 0;
}

int stream_push(y3g_stream* s, int slot, unsigned gen, const y3g_integral* igs, size_t n,
                unsigned long long* first)
{
  if (!s || (!igs && n) || !first || slot < 0 || slot >= s->cfg.table_slots)
    return fail(Y3G_ERR_ARG, "stream_push: bad arguments");
  if (stream_room(s) < n) return fail(Y3G_ERR_STATE, "stream_push: the ring is full");
  for (size_t i = 0; i < n; ++i) {
    y3g_integral const& g = igs[i];
    if (g.ndim < 2 || g.ndim > kMaxDim || g.kind < 0 || g.kind > Y3G_DSIGMA6D ||
        g.algorithm != Y3G_PAGANI || !(g.eps_rel >= 0) || g.max_eval < 1)
      return fail(Y3G_ERR_ARG, "stream_push: integral " + std::to_string(i) +
                                 " is not a cubature integral of kinds 0-2");
    for (int a = 0; a < g.ndim; ++a)
      if (!(g.highs[a] > g.lows[a]))
        return fail(Y3G_ERR_ARG, "stream_push: empty volume in integral " + std::to_string(i));
  }
  *first = s->tail;
  for (size_t i = 0; i < n; ++i) {
    size_t const k = (size_t)((s->tail + i) % (unsigned long long)s->cfg.ring_items);
    StreamItem e;
    std::memset(&e, 0, sizeof(e));
    e.job = kStreamJobIntegral;
    e.slot = slot;
    e.gen = gen;
    e.ig = igs[i];
    e.ig.point = slot;
    s->ring[k] = e;
    s->kind[k] = 1;
  }
  stream_publish_tail(s, n);
  return 0;
}

size_t stream_collect(y3g_stream* s, void (*f)(void*, unsigned long long, const y3g_result&),
                      void* ctx)
{
  size_t got = 0;
  unsigned long long const cap = (unsigned long long)s->cfg.ring_items;
  for (unsigned long long i = s->low; i < s->tail; ++i) {
    size_t const k = (size_t)(i % cap);
    if (!s->kind[k]) continue;  // reported already
    if (__atomic_load_n(&s->seq[k], __ATOMIC_ACQUIRE) != i + 1) continue;
    y3g_result const r = s->res[k];
    if (s->kind[k] == 2) {
      s->tables_pending[s->ring[k].slot] -= 1;
    } else {
      f(ctx, i, r);
      ++got;
    }
    s->kind[k] = 0;
  }
  while (s->low < s->tail && !s->kind[(size_t)(s->low % cap)]) ++s->low;
  return got;
}

int stream_kick(y3g_stream* s)
{
  if (s->running) {
    hipError_t const q = hipEventQuery(s->done);
    if (q == hipErrorNotReady) return 0;
    if (q != hipSuccess) return fail(Y3G_ERR_HIP, std::string("stream launch: ") +
                                                    hipGetErrorString(q));
    s->running = false;
    HIPCHK(hipMemcpy(&s->head_known, &s->dev->head, sizeof(unsigned long long),
                     hipMemcpyDeviceToHost));
  }
  if (s->tail <= s->head_known) return 0;
  HIPCHK(hipSetDevice(s->rt->device));
  // a new launch starts with no busy workgroup and without the quit flag
  HIPCHK(hipMemsetAsync(&s->dev->busy, 0, sizeof(unsigned), s->rt->stream));
  HIPCHK(hipMemsetAsync(&s->dev->quit, 0, sizeof(unsigned), s->rt->stream));
  HIPCHK(launch_cubature_stream(s->A, s->Z, s->cfg.options.precision, s->workgroups,
                                s->rt->stream));
  HIPCHK(hipEventRecord(s->done, s->rt->stream));
  s->running = true;
  s->launches += 1;
  return 0;
}

long long stream_launches(const y3g_stream* s) { return s->launches; }

int stream_close(y3g_stream* s, long long* nevals)
{
  if (!s) return 0;
  __atomic_store_n(&s->ctl->close, 1u, __ATOMIC_RELEASE);
  int rc = 0;
  if (s->running && hipStreamSynchronize(s->rt->stream) != hipSuccess)
    rc = fail(Y3G_ERR_HIP, "stream_close: the launch failed");
  s->running = false;
  unsigned long long nev = 0;
  if (hipMemcpy(&nev, s->rt->d_nevals, sizeof(nev), hipMemcpyDeviceToHost) != hipSuccess &&
      !rc)
    rc = fail(Y3G_ERR_HIP, "stream_close: counters");
  if (nevals) *nevals = (long long)nev;
  stream_free(s);
  return rc;
}

}  // namespace y3g

// ------------------------------------------------------------- HIP graphs
//
// y3g_graph_*: one batched step captured once on the runtime's stream --
// the points' constants in (pinned), k_tables_build, the descriptors and the
// longest-first order in (pinned), the queue and counters cleared by a fill
// kernel (no memset node), k_cubature, results and the evaluation count out
// (pinned) -- and replayed by hipGraphLaunch between two event records on
// the stream (the step's time: the launch and its copies and table build).

struct y3g_graph {
  y3g_runtime* rt = nullptr;
  y3g_table_spec spec{};
  y3g_options opt{};
  int n_points = 0;
  size_t n = 0;
  int slots = 0;
  unsigned slab_gen = 0;
  // the caller's inputs (host) and the replay's pinned staging
  std::vector<y3g_params> params;
  std::vector<y3g_integral> igs;
  PointDev* h_points = nullptr;
  y3g_integral* h_igs = nullptr;
  int* h_order = nullptr;
  y3g_result* h_res = nullptr;
  unsigned long long* h_nev = nullptr;
  // device buffers
  PointDev* d_points = nullptr;
  double *hmf64 = nullptr, *da64 = nullptr, *sci64 = nullptr;
  float *hmf32 = nullptr, *da32 = nullptr, *sci32 = nullptr;
  double* hmfc64 = nullptr;
  float* hmfc32 = nullptr;
  y3g_integral* d_igs = nullptr;
  y3g_result* d_res = nullptr;
  int* d_order = nullptr;
  unsigned* d_counters = nullptr;  // queue (int), pad, evaluation count (u64)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static void graph_free(y3g_graph* g)
{
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->h_points) (void)hipHostFree(g->h_points);
  if (g->h_igs) (void)hipHostFree(g->h_igs);
  if (g->h_order) (void)hipHostFree(g->h_order);
  if (g->h_res) (void)hipHostFree(g->h_res);
  if (g->h_nev) (void)hipHostFree(g->h_nev);
  dfree(g->d_points);
  dfree(g->hmf64), dfree(g->da64), dfree(g->sci64);
  dfree(g->hmf32), dfree(g->da32), dfree(g->sci32);
  dfree(g->hmfc64), dfree(g->hmfc32);
  dfree(g->d_igs), dfree(g->d_res), dfree(g->d_order), dfree(g->d_counters);
  delete g;
}

extern "C" {

int y3g_graph_create(y3g_runtime* rt, const y3g_table_spec* spec, int n_points, size_t n,
                     const y3g_options* opt, y3g_graph** out)
{
  if (!rt || !spec || !out || n_points <= 0 || n == 0 || n > (size_t)(1 << 30))
    return fail(Y3G_ERR_ARG, "y3g_graph_create: bad arguments");
  *out = nullptr;
  if (int rc = check_spec(spec)) return rc;
  y3g_options o;
  y3g_options_default(&o);
  if (opt) o = *opt;
  if (o.single || o.max_regions < 64 || o.max_split_per_step < 1 ||
      o.max_split_per_step > 1024 || o.n_bands < 0 || o.n_bands > 8 ||
      o.ext_regions < 0 || o.ext_regions > o.max_regions || !(o.ext_ratio >= 0.0) ||
      (o.precision != Y3G_FP32SPLIT && o.precision != Y3G_FP64))
    return fail(Y3G_ERR_ARG, "y3g_graph_create: options out of range (batched cubature, "
                             "fp32split or fp64)");
  for (int k = 0; k < o.n_bands; ++k)
    if (o.band_regions[k] < 64 || o.band_regions[k] > o.max_regions ||
        (k > 0 && !(o.band_eps[k] > o.band_eps[k - 1])))
      return fail(Y3G_ERR_ARG, "y3g_graph_create: tolerance bands need ascending edges and "
                               "budgets in [64, max_regions]");
  HIPCHK(hipSetDevice(rt->device));
  auto* g = new y3g_graph;
  auto bail = [&](int rc) {
    graph_free(g);
    return rc;
  };
  g->rt = rt;
  g->spec = *spec;
  g->opt = o;
  g->n_points = n_points;
  g->n = n;
  g->params.assign(n_points, y3g_params{});
  for (auto& p : g->params) y3g_default_params(&p);
  g->igs.assign(n, y3g_integral{});
  int const block = o.block_threads == 256 ? 256 : 512;
  g->slots = std::max(1, std::min<int>(resident_slots(rt, o, Y3G_PAGANI), (int)n));
  if (int rc = ensure_slab(rt, std::max(g->slots, rt->slab_slots), o.max_regions))
    return bail(rc);
  g->slab_gen = rt->slab_gen;
  size_t const hs = (size_t)spec->n_lnm * spec->n_z;
  size_t const hcs = (size_t)hmfc_stride_of(*spec);
  auto pin = [](auto** p, size_t cnt) {
    return hipHostMalloc(reinterpret_cast<void**>(p), sizeof(**p) * cnt, hipHostMallocDefault);
  };
  if (pin(&g->h_points, n_points) != hipSuccess || pin(&g->h_igs, n) != hipSuccess ||
      pin(&g->h_order, n) != hipSuccess || pin(&g->h_res, n) != hipSuccess ||
      pin(&g->h_nev, 1) != hipSuccess)
    return bail(fail(Y3G_ERR_MEMORY, "y3g_graph_create: pinned host memory"));
  if (dalloc(&g->d_points, n_points) || dalloc(&g->hmf64, n_points * hs) ||
      dalloc(&g->hmf32, n_points * hs) || dalloc(&g->da64, (size_t)n_points * spec->n_da) ||
      dalloc(&g->da32, (size_t)n_points * spec->n_da) ||
      dalloc(&g->sci64, (size_t)n_points * spec->n_sci) ||
      dalloc(&g->sci32, (size_t)n_points * spec->n_sci) || dalloc(&g->d_igs, n) ||
      dalloc(&g->d_res, n) || dalloc(&g->d_order, n) || dalloc(&g->d_counters, 4) ||
      dalloc(&g->hmfc64, n_points * hcs) || dalloc(&g->hmfc32, n_points * hcs))
    return bail(Y3G_ERR_MEMORY);
  TableArgs T;
  T.points = g->d_points;
  T.spec = *spec;
  T.hmf64 = g->hmf64, T.hmf32 = g->hmf32, T.hmf_stride = (int)hs;
  T.da64 = g->da64, T.da32 = g->da32, T.da_stride = spec->n_da;
  T.sci64 = g->sci64, T.sci32 = g->sci32, T.sci_stride = spec->n_sci;
  T.hmfc64 = g->hmfc64, T.hmfc32 = g->hmfc32, T.hmfc_stride = (int)hcs;
  KernArgs A;
  std::memset(&A, 0, sizeof(A));
  A.igs = g->d_igs;
  A.n = (int)n;
  A.order = g->d_order;
  A.points = g->d_points;
  A.hmf32 = g->hmf32, A.hmf64 = g->hmf64, A.hmf_stride = (int)hs;
  A.da32 = g->da32, A.da64 = g->da64, A.da_stride = spec->n_da;
  A.sci32 = g->sci32, A.sci64 = g->sci64, A.sci_stride = spec->n_sci;
  if (o.hmf_smooth && smooth_ok(*spec))
    A.hmfc32 = g->hmfc32, A.hmfc64 = g->hmfc64, A.hmfc_stride = (int)hcs;
  A.slab_lo = rt->slab_lo, A.slab_hi = rt->slab_hi;
  A.slab_est = rt->slab_est, A.slab_err = rt->slab_err, A.slab_raw = rt->slab_raw;
  A.slab_ax = rt->slab_ax;
  A.cap = o.max_regions;
  A.kmax = o.max_split_per_step;
  A.init_split = o.init_split;
  A.init_lnm = o.init_split_lnm;
  A.err_scale = o.err_scale;
  A.ext_ratio = o.ext_ratio;
  A.ext_regions = o.ext_regions;
  A.err_c = o.err_c, A.err_smin = o.err_smin;
  A.tl_c = o.tl_c, A.tl_smin = o.tl_smin;
  A.queue = reinterpret_cast<int*>(g->d_counters);
  A.nevals = reinterpret_cast<unsigned long long*>(g->d_counters + 2);
  A.out = g->d_res;
  hipStream_t const st = rt->stream;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return bail(fail(Y3G_ERR_HIP, "y3g_graph_create: stream capture"));
  bool ok = hipMemcpyAsync(g->d_points, g->h_points, sizeof(PointDev) * n_points,
                           hipMemcpyHostToDevice, st) == hipSuccess &&
            launch_tables(T, n_points, st) == hipSuccess &&
            hipMemcpyAsync(g->d_igs, g->h_igs, sizeof(y3g_integral) * n, hipMemcpyHostToDevice,
                           st) == hipSuccess &&
            hipMemcpyAsync(g->d_order, g->h_order, sizeof(int) * n, hipMemcpyHostToDevice,
                           st) == hipSuccess &&
            launch_zero_words(g->d_counters, 4, st) == hipSuccess &&
            launch_cubature(A, o.precision, block, g->slots, st) == hipSuccess &&
            hipMemcpyAsync(g->h_res, g->d_res, sizeof(y3g_result) * n, hipMemcpyDeviceToHost,
                           st) == hipSuccess &&
            hipMemcpyAsync(g->h_nev, A.nevals, sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, st) == hipSuccess;
  hipError_t const ec = hipStreamEndCapture(st, &g->graph);
  if (!ok || ec != hipSuccess)
    return bail(fail(Y3G_ERR_HIP, "y3g_graph_create: capture failed"));
  if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
    return bail(fail(Y3G_ERR_HIP, "y3g_graph_create: instantiation failed"));
  *out = g;
  return 0;
}

y3g_params* y3g_graph_params(y3g_graph* g) { return g ? g->params.data() : nullptr; }
y3g_integral* y3g_graph_integrals(y3g_graph* g) { return g ? g->igs.data() : nullptr; }
const y3g_result* y3g_graph_results(const y3g_graph* g) { return g ? g->h_res : nullptr; }

int y3g_graph_run(y3g_graph* g)
{
  if (!g) return fail(Y3G_ERR_ARG, "y3g_graph_run: null graph");
  y3g_runtime* rt = g->rt;
  if (rt->slab_gen != g->slab_gen)
    return fail(Y3G_ERR_STATE, "y3g_graph_run: the runtime's region slab was reallocated "
                               "since the graph was captured");
  y3g_options const& o = g->opt;
  for (size_t i = 0; i < g->n; ++i) {
    y3g_integral const& ig = g->igs[i];
    if (ig.ndim < 2 || ig.ndim > kMaxDim || ig.kind < 0 || ig.kind > Y3G_DSIGMA6D ||
        ig.point < 0 || ig.point >= g->n_points || ig.algorithm != Y3G_PAGANI ||
        !(ig.eps_rel >= 0) || ig.max_eval < 1)
      return fail(Y3G_ERR_ARG, "y3g_graph_run: integral " + std::to_string(i) +
                                 " is malformed (cubature, kinds 0-2, point in range)");
    for (int a = 0; a < ig.ndim; ++a)
      if (!(ig.highs[a] > ig.lows[a]))
        return fail(Y3G_ERR_ARG, "y3g_graph_run: empty volume in integral " +
                                   std::to_string(i));
  }
  for (int k = 0; k < g->n_points; ++k) g->h_points[k] = make_point(g->params[k], g->spec);
  std::vector<int> ids(g->n);
  std::iota(ids.begin(), ids.end(), 0);
  for (size_t i = 0; i < g->n; ++i) {
    y3g_integral v = g->igs[i];
    // the tolerance bands' region budgets, as y3g_integrate applies them
    for (int k = 0; k < o.n_bands; ++k)
      if (v.eps_rel <= o.band_eps[k]) {
        long long const lim = 2LL * o.band_regions[k] * gm_num_nodes(v.ndim);
        if (lim < v.max_eval) v.max_eval = lim;
        break;
      }
    g->h_igs[i] = v;
  }
  std::vector<int> order;
  lpt_order(g->h_igs, ids, order);
  std::memcpy(g->h_order, order.data(), sizeof(int) * g->n);
  HIPCHK(hipSetDevice(rt->device));
  HIPCHK(hipEventRecord(rt->ev0, rt->stream));
  HIPCHK(hipGraphLaunch(g->exec, rt->stream));
  HIPCHK(hipEventRecord(rt->ev1, rt->stream));
  HIPCHK(hipStreamSynchronize(rt->stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, rt->ev0, rt->ev1));
  rt->last_ms = ms;
  rt->last_nevals = (long long)*g->h_nev;
  return 0;
}

int y3g_graph_destroy(y3g_graph* g)
{
  if (!g) return 0;
  (void)hipStreamSynchronize(g->rt->stream);
  graph_free(g);
  return 0;
}

}  // extern "C"
