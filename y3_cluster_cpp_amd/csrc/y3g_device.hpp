// y3g_device.hpp -- device model functors and the three hot-path integrands
// for gfx950, templated on the compute precision T (float for the FP32
// integrand of the split path, double for the plain FP64 path).
//
// Functor call signatures follow the reference functors:
//   hmf(lnM, zt)                   src/models/hmf_t.cuh:54-60
//   mor(lt, lnM, zt)               src/models/mor_t.cuh:70-81 (lognormal form)
//   lc_lt(lc, lt, zt)              src/models/lc_lt_t.cuh:36-68 (fmsk = 0)
//   lo_lc(lo, lc, R)               src/models/lo_lc_t.cuh:20-30
//   int_zo_zt(zomin, zomax, zt)    src/models/int_zo_zt_des_t.cuh:11-17
//   dv_do_dz(zt)                   src/models/dv_do_dz_t.cuh:19-27
//   roffset(x)                     src/models/roffset_t.cuh:15-21
//   nfw_dsigma_mis(r, rmis, lnM)   src/models/nfw_dsigma_mis.cuh:104-123
// Integrand bodies: NC_cent_y1_mort_cuda_scalarintegrand.cu:164-175 (3-D),
// numberCounts.cu:159-174 (5-D), sigma_miscent_y1_mort_cuda_scalarintegrand.cu
// :161-182 (6-D).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "y3g.h"

namespace y3g {

constexpr int kMaxDim = 6;
constexpr int kMaxHmf = 200 * 32;
constexpr int kMaxDa = 512;
constexpr int kMaxSci = 128;
constexpr double kLnPivotHmf = 32.236191301916641;  // ln 1e14
constexpr double kLnPivotMor = 33.334803590584751;  // ln 3e14
constexpr double kRhoCrit = 2.77533742639e+11;      // h^2 Msun / Mpc^3
constexpr double kDvConst = 2997.92;                // dv_do_dz_t.cuh:25
constexpr double kLn3e14Over800Pi = 25.505461977067423;  // ln(3e14 / (800 pi))
constexpr double kDh = 2997.92458;                  // c / H0 [Mpc/h]
// 4 pi G / c^2 in pc^2/Msun per Mpc of distance, from G = 4.51710305e-48
// Mpc^3/Msun/s^2 and c = 9.71561189e-15 Mpc/s
// (src/modules/sigma_park_y1/buildSigmaCritInv.py:56-57)
constexpr double kSciConst = 4.0 * 3.14159265358979323846 * 4.51710305e-48 /
                             (9.71561189e-15 * 9.71561189e-15) * 1e12;
constexpr double kTMax = 7.0;
constexpr double kTSoft = 4.0;  // sharpness of the t window's smoothed clip

// Per-parameter-point constants (host-prepared).
struct PointDev {
  y3g_params p;
  double lnm0, dlnm, inv_dlnm;
  double z0, dz, inv_dz;
  double inv_dzda, inv_dzsci;
  double rho_m;
  int nm, nz, nda, nsci;
};

// ---------------------------------------------------------------- math
template <typename T> struct M;
// FP32: the hardware transcendental instructions (v_exp_f32, v_log_f32,
// v_sqrt_f32, v_cos_f32; about 1 ulp on the arguments met here) in place of
// the correctly-rounded library sequences -- the FP32 integrand's inputs
// already carry errors of that size.  The logarithm is v_log_f32 (log2)
// times ln 2, without the denormal-input scaling of the library form (no
// argument here is denormal).  FP64 keeps the library functions.
template <> struct M<float> {
  static __device__ __forceinline__ float exp(float x) { return __expf(x); }
  static __device__ __forceinline__ float log(float x)
  {
    return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
  }
  // log1p through the hardware log: log(1 + x) x / ((1 + x) - 1), which
  // cancels the rounding of 1 + x (D. Goldberg, "What every computer
  // scientist should know about floating-point arithmetic", 1991, Thm 4)
  static __device__ __forceinline__ float log1p(float x)
  {
    float const p = 1.0f + x;
    return p == 1.0f ? x : log(p) * x * __builtin_amdgcn_rcpf(p - 1.0f);
  }
  static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float erf(float x) { return erff(x); }
  static __device__ __forceinline__ float erfc(float x) { return erfcf(x); }
  static __device__ __forceinline__ float erfcx(float x) { return erfcxf(x); }
  static __device__ __forceinline__ float cos(float x) { return __cosf(x); }
  static __device__ __forceinline__ float atan(float x) { return atanf(x); }
  static __device__ __forceinline__ float cbrt(float x) { return cbrtf(x); }
  static __device__ __forceinline__ float fabs(float x) { return fabsf(x); }
  static __device__ __forceinline__ float pow(float x, float y) { return powf(x, y); }
  // a / b as a times the hardware reciprocal (1 ulp), a / sqrt(2)
  static __device__ __forceinline__ float div(float a, float b)
  {
    return a * __builtin_amdgcn_rcpf(b);
  }
  static __device__ __forceinline__ float div_sqrt2(float a) { return a * 0.70710678118654752f; }
};
template <> struct M<double> {
  static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
  static __device__ __forceinline__ double log(double x) { return ::log(x); }
  static __device__ __forceinline__ double log1p(double x) { return ::log1p(x); }
  static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
  static __device__ __forceinline__ double erf(double x) { return ::erf(x); }
  static __device__ __forceinline__ double erfc(double x) { return ::erfc(x); }
  static __device__ __forceinline__ double erfcx(double x) { return ::erfcx(x); }
  static __device__ __forceinline__ double cos(double x) { return ::cos(x); }
  static __device__ __forceinline__ double atan(double x) { return ::atan(x); }
  static __device__ __forceinline__ double cbrt(double x) { return ::cbrt(x); }
  static __device__ __forceinline__ double fabs(double x) { return ::fabs(x); }
  static __device__ __forceinline__ double pow(double x, double y) { return ::pow(x, y); }
  static __device__ __forceinline__ double div(double a, double b) { return a / b; }
  static __device__ __forceinline__ double div_sqrt2(double a) { return a / 1.4142135623730951; }
};

#define Y3G_HD __device__ __forceinline__

// Tables of one point as seen by an integrand evaluation (LDS or global).
template <typename T> struct Tabs {
  const T* hmf;
  const T* da;
  const T* sci;
  // the HMF table's cubic-spline coefficients (hmf_coefs): when set, the
  // synthetic integrands read the hat-smoothed HMF (hmf_smooth) instead of
  // the bilinear one
  const T* hmfc = nullptr;
};

// ------------------------------------------------------------ functors
// primitives.hh:60-65
template <typename T> Y3G_HD T gaussian(T x, T mu, T sigma)
{
  T const z = M<T>::div(x - mu, sigma);
  return M<T>::div(M<T>::exp(T(-0.5) * z * z) * T(0.3989422804014327), sigma);
}

// roffset_t.cuh:15-21 (Gamma k = 2 kernel of the reference).
template <typename T> Y3G_HD T roffset_gamma(T x, T tau)
{
  return x / tau / tau * M<T>::exp(-x / tau);
}

// lo_lc_t.cuh:20-30
template <typename T>
Y3G_HD T lo_lc(T lo, T lc, T R, T alpha, T a, T b, T R_lambda)
{
  T const x = R / R_lambda;
  T const y = lo / lc;
  T const mu_y = M<T>::exp(-x * x / alpha / alpha);
  T const sigma_y = a * M<T>::atan(b * x);
  return gaussian<T>(y, mu_y, sigma_y) / lc;
}

// sigma_photoz_des.cuh:7-36: Horner in (clamp(z, 0.15, 0.7) - 0.4).
template <typename T> Y3G_HD T sigma_photoz_des(T zt)
{
  const T c[11] = {T(-40358.8315),  T(2798.08304),   T(9333.80185),
                   T(-657.348248),  T(-840.565610),  T(46.8506649),
                   T(37.8839498),   T(-0.868811858), T(-0.808928182),
                   T(0.00890199353), T(0.0139811265)};
  T z = zt < T(0.15) ? T(0.15) : (zt > T(0.7) ? T(0.7) : zt);
  z -= T(0.4);
  T s = T(0);
  for (int i = 0; i < 10; ++i) s = (c[i] + s) * z;
  return s + c[10];
}

__device__ __forceinline__ float erfcx_pos_f(float x);

// Half the difference erf(b) - erf(a), b >= a, in the cancellation-free
// form (complementary functions on the tails).  FP32: erfc(|x|) =
// exp(-x^2) erfcx(|x|) with the polynomial erfcx below, branch-free.
template <typename T> Y3G_HD T half_erf_diff(T a, T b)
{
  if constexpr (sizeof(T) == 4) {
    float const aa = fabsf(a), ab = fabsf(b);
    float const ea = __expf(-aa * aa) * erfcx_pos_f(aa);
    float const eb = __expf(-ab * ab) * erfcx_pos_f(ab);
    return a > 0.0f ? 0.5f * (ea - eb) : (b < 0.0f ? 0.5f * (eb - ea) : 0.5f * (2.0f - ea - eb));
  }
  if (a > T(0)) return T(0.5) * (M<T>::erfc(a) - M<T>::erfc(b));
  if (b < T(0)) return T(0.5) * (M<T>::erfc(-b) - M<T>::erfc(-a));
  return T(0.5) * (M<T>::erf(b) - M<T>::erf(a));
}

// lo_lc integrated over a lambda_obs bin [lo_lo, lo_hi] in closed form: with
// y = lo / lc ~ N(mu_y, sigma_y) and dy = dlo / lc,
//   int lo_lc dlo = Phi((lo_hi / lc - mu_y) / sigma_y) - Phi((lo_lo / lc - mu_y) / sigma_y);
// sigma_y = 0 (R = 0) leaves the indicator of lo_lo <= mu_y lc < lo_hi.
template <typename T>
Y3G_HD T lo_lc_bin(T lo_lo, T lo_hi, T lc, T R, T alpha, T a, T b, T R_lambda)
{
  T const x = R / R_lambda;
  T const mu_y = M<T>::exp(-x * x / alpha / alpha);
  T const sigma_y = a * M<T>::atan(b * x);
  if (!(sigma_y > T(0))) return (lo_lo <= mu_y * lc && mu_y * lc < lo_hi) ? T(1) : T(0);
  T const s = T(1.4142135623730951) * sigma_y;
  return half_erf_diff<T>((lo_lo / lc - mu_y) / s, (lo_hi / lc - mu_y) / s);
}

// int_zo_zt_des_t.cuh:11-17 with the DES photo-z polynomial.
template <typename T> Y3G_HD T int_zo_zt_des(T zomin, T zomax, T zt)
{
  T const base = T(1.4142135623730951) * sigma_photoz_des<T>(zt);
  return half_erf_diff<T>(M<T>::div(zomin - zt, base), M<T>::div(zomax - zt, base));
}

// z_obs ~ N(zt, sigz0 (1+zt)): pdf and bin integral.
template <typename T> Y3G_HD T zo_pdf(T zo, T zt, T sigz0)
{
  return gaussian<T>(zo, zt, sigz0 * (T(1) + zt));
}
template <typename T> Y3G_HD T zo_bin(T zlo, T zhi, T zt, T sigz0)
{
  T const base = T(1.4142135623730951) * sigz0 * (T(1) + zt);
  return half_erf_diff<T>(M<T>::div(zlo - zt, base), M<T>::div(zhi - zt, base));
}

// FP32 erfcx on x >= 0 by the form of M. M. Shepherd and J. G. Laframboise,
// "Chebyshev approximation of (1 + 2x) exp(x^2) erfc x in 0 <= x < infinity",
// Math. Comp. 36 (1981): (1 + 2x) erfcx(x) as a degree-10 polynomial in
// t = (x - 2) / (x + 2), coefficients fitted by csrc/fit_erfcx.py (largest
// relative error 2.7e-7 in FP32 on [0, 1e7]); branch-free, two reciprocals.
__device__ __forceinline__ float erfcx_pos_f(float x)
{
  float const t = (x - 2.0f) * __builtin_amdgcn_rcpf(x + 2.0f);
  float p = 0.00016080358f;
  p = p * t - 0.00018365444f;
  p = p * t - 0.0013957069f;
  p = p * t + 0.0010431446f;
  p = p * t + 0.0086891595f;
  p = p * t - 0.007942721f;
  p = p * t - 0.05421196f;
  p = p * t + 0.16403584f;
  p = p * t - 0.16603109f;
  p = p * t - 0.09276308f;
  p = p * t + 1.2769784f;
  return p * __builtin_amdgcn_rcpf(1.0f + 2.0f * x);
}

// FP32 line-of-sight pdf, Gaussian (sigma^2 = lt) plus the fraction fprj in
// the EMG tail with tau = ips / lt (ips = 1 / prj_scale): one reciprocal
// square root gives 1 / sigma, one exponential serves both terms, and
// tau sigma^2 = ips.  Same quantity as lo_lt_pdf below.
__device__ __forceinline__ float lo_lt_pdf_f(float lo, float lt, float fprj, float ips)
{
  float const isig = __builtin_amdgcn_rsqf(lt);
  float const z = (lo - lt) * isig;
  float const e = __expf(-0.5f * z * z);
  float const g = e * 0.3989422804014327f * isig;
  if (fprj == 0.0f) return g;
  float const B = (lt + ips - lo) * isig * 0.70710678118654752f;
  // E = exp(-z^2/2) erfcx(|B|); the pdf is tau/2 E for B >= 0 and
  // tau/2 (2 exp(A) - E) = tau/2 exp(A) erfc(B) for B < 0
  float const E = e * erfcx_pos_f(fabsf(B));
  float const tau = ips * isig * isig;
  float emg;
  if (B >= 0.0f) {
    emg = 0.5f * tau * E;
  } else {
    float const A = tau * (lt - lo) + 0.5f * tau * ips;
    emg = 0.5f * tau * (2.0f * __expf(A) - E);
  }
  return (1.0f - fprj) * g + fprj * emg;
}

// Exponentially-modified Gaussian pdf (lc_lt_t.cuh:56-61 second term with
// fmsk = 0; emg_des_t.cuh:172-193), overflow-free through erfcx.
template <typename T> Y3G_HD T emg_pdf(T x, T mu, T sigma, T tau)
{
  T const B = M<T>::div(mu + tau * sigma * sigma - x, T(1.4142135623730951) * sigma);
  if (B >= T(0)) {
    T const d = M<T>::div(x - mu, sigma);
    return T(0.5) * tau * M<T>::exp(T(-0.5) * d * d) * M<T>::erfcx(B);
  }
  T const A = tau * (mu - x) + T(0.5) * tau * tau * sigma * sigma;
  return T(0.5) * tau * M<T>::exp(A) * M<T>::erfc(B);
}

// EMG CDF (emg_des_t.cuh:136-168).
template <typename T> Y3G_HD T emg_cdf(T x, T mu, T sigma, T tau)
{
  T const z = M<T>::div(x - mu, sigma);
  T const phi = T(0.5) * M<T>::erfc(M<T>::div_sqrt2(-z));
  T const u = M<T>::div_sqrt2(tau * sigma - z);
  T tail;
  if (u >= T(0))
    tail = T(0.5) * M<T>::exp(T(-0.5) * z * z) * M<T>::erfcx(u);
  else
    tail = T(0.5) *
           M<T>::exp(-tau * (x - mu) + T(0.5) * tau * tau * sigma * sigma) *
           M<T>::erfc(u);
  return phi - tail;
}

// 1 - F_EMG(x) = Phi(-z) + exp(A) Phi(z - tau sigma).
template <typename T> Y3G_HD T emg_sf(T x, T mu, T sigma, T tau)
{
  T const z = M<T>::div(x - mu, sigma);
  T const q = T(0.5) * M<T>::erfc(M<T>::div_sqrt2(z));
  T const u = M<T>::div_sqrt2(tau * sigma - z);
  T tail;
  if (u >= T(0))
    tail = T(0.5) * M<T>::exp(T(-0.5) * z * z) * M<T>::erfcx(u);
  else
    tail = T(0.5) *
           M<T>::exp(-tau * (x - mu) + T(0.5) * tau * tau * sigma * sigma) *
           M<T>::erfc(u);
  return q + tail;
}

// F(hi) - F(lo) of the EMG, from the survival function when F(lo) > 1/2.
template <typename T> Y3G_HD T emg_cdf_diff(T lo, T hi, T mu, T sigma, T tau)
{
  T const flo = emg_cdf<T>(lo, mu, sigma, tau);
  if (flo > T(0.5)) return emg_sf<T>(lo, mu, sigma, tau) - emg_sf<T>(hi, mu, sigma, tau);
  return emg_cdf<T>(hi, mu, sigma, tau) - flo;
}

// P(lo | lt) = (1 - f_prj) N(lt, sqrt lt) + f_prj EMG(lt, sqrt lt, 1/(s lt)).
template <typename T> Y3G_HD T lo_lt_pdf(T lo, T lt, T fprj, T pscale)
{
  if constexpr (sizeof(T) == 4) return lo_lt_pdf_f(lo, lt, fprj, __builtin_amdgcn_rcpf(pscale));
  T const sigma = M<T>::sqrt(lt);
  T const g = gaussian<T>(lo, lt, sigma);
  if (fprj == T(0)) return g;
  T const tau = M<T>::div(T(1), pscale * lt);
  return (T(1) - fprj) * g + fprj * emg_pdf<T>(lo, lt, sigma, tau);
}

// Integral of lo_lt_pdf over [lo_lo, lo_hi].
template <typename T> Y3G_HD T lo_lt_bin(T lo_lo, T lo_hi, T lt, T fprj, T pscale)
{
  T const sigma = M<T>::sqrt(lt);
  T const base = T(1.4142135623730951) * sigma;
  T const g = half_erf_diff<T>(M<T>::div(lo_lo - lt, base), M<T>::div(lo_hi - lt, base));
  if (fprj == T(0)) return g;
  T const tau = M<T>::div(T(1), pscale * lt);
  T const e = emg_cdf_diff<T>(lo_lo, lo_hi, lt, sigma, tau);
  return (T(1) - fprj) * g + fprj * e;
}

// Uniform-grid linear interpolation with clamp (interp_1d.hh clamp()).
template <typename T>
Y3G_HD T lerp_clamp(const T* ys, int n, double x_off, double inv_dx)
{
  double u = x_off * inv_dx;
  u = u < 0.0 ? 0.0 : (u > double(n - 1) ? double(n - 1) : u);
  int i = int(u);
  if (i > n - 2) i = n - 2;
  T const t = T(u - double(i));
  return ys[i] + t * (ys[i + 1] - ys[i]);
}

// HMF_t::operator() (hmf_t.cuh:54-60): clamped bilinear lookup of the
// (n_z, n_lnM) table, x-fastest, times s (lnM log10 e - 13.81) + q.
template <typename T>
Y3G_HD T hmf(const PointDev& P, const T* tab, double lnM, double zt)
{
  double u = (lnM - P.lnm0) * P.inv_dlnm;
  double v = (zt - P.z0) * P.inv_dz;
  u = u < 0.0 ? 0.0 : (u > double(P.nm - 1) ? double(P.nm - 1) : u);
  v = v < 0.0 ? 0.0 : (v > double(P.nz - 1) ? double(P.nz - 1) : v);
  int i = int(u), j = int(v);
  if (i > P.nm - 2) i = P.nm - 2;
  if (j > P.nz - 2) j = P.nz - 2;
  T const t = T(u - double(i)), s = T(v - double(j));
  const T* r0 = tab + j * P.nm + i;
  const T* r1 = r0 + P.nm;
  T const a = r0[0] + t * (r0[1] - r0[0]);
  T const b = r1[0] + t * (r1[1] - r1[0]);
  return (a + s * (b - a)) *
         T(P.p.hmf_s * (lnM * 0.4342944819 - 13.8124426028) + P.p.hmf_q);
}

// ------------------------------------------------------ hat-smoothed HMF
// The bilinear HMF is kinked along every knot line, and a cubature rule
// straddling the kinks sees a sawtooth of relative size h^2 f''/8 (up to
// 5e-3 along z at high mass): errors that are aliasing rather than
// truncation.  Smoothing identity: with S any interpolant of the table
// values and k_h the hat kernel of one cell (unit mass),
//   Int H g = h sum_i S(x_i) (g * k_h)(x_i)   (exact, H the bilinear form)
//   Int (S * k_h) g = Int S (g * k_h),
// which differ by the error of a knot sum of the smooth S (g * k_h) against
// its integral: by Poisson summation the Fourier transform of that product
// at multiples of 2 pi / h, where k_h's transform sinc^2(k h / 2) vanishes.
// With S the cubic spline sum_k c_k B3((x - x_k)/h), S * k_h =
// sum_k c_k B5((x - x_k)/h) (B5 the centred quintic B-spline): a C4 surface
// whose integral against the integrands' other factors equals that of the
// bilinear table to 1e-9 .. 1e-7 relative (tests/test_oracle_smooth.py), and
// on which the Genz-Malik rules converge as on a smooth function.  The
// oracle restates it (y3o_hmf_coefs, y3o_hmf_smooth).
//
// Coefficients, per axis: the not-a-knot cubic spline (spline_coefs), the
// (1, 4, 1) system by Thomas elimination with pivots beta_1 = 4, beta_i =
// 4 - 1 / beta_{i-1}, held at beta_32 beyond i = 32 (converged to double
// precision).  Layout [(l + 2) (n_lnm + 4) + k + 2], l in [-2, n_z + 1], k in
// [-2, n_lnm + 1]; tables with fewer than kSmoothMinKnots knots on an axis
// are read bilinearly.
constexpr int kSmoothMinKnots = 6;
constexpr int kSplineBetaCap = 32;
__host__ __device__ inline double spline_beta(int i)
{
  double b = 4.0;
  int const m = i < kSplineBetaCap ? i : kSplineBetaCap;
  for (int k = 2; k <= m; ++k) b = 4.0 - 1.0 / b;
  return b;
}
// y[i * ys], i < n (n >= 6) -> c[(k + 2) * cs], k in [-2, n + 1]; y may
// alias c with the same stride (y = c + 2 cs).  Not-a-knot ends: one cubic
// over [x_0, x_2] (and [x_{n-3}, x_{n-1}]) fixes c_1 = y_1 - (y_0 - 2 y_1 +
// y_2) / 6 (S'' at x_1 is the central second difference); the (1, 4, 1)
// system over c_2 .. c_{n-3} by Thomas elimination; then c_0, c_{-1} (and
// c_{n-1}, c_n) from the interpolation rows, and one more coefficient per end
// continuing the end cubic.
__host__ __device__ inline void spline_coefs(const double* y, int n, long ys, double* c, long cs)
{
  auto C = [&](int k) -> double& { return c[(long)(k + 2) * cs]; };
  double const y0 = y[0], y1 = y[ys], y2 = y[2 * ys];
  double const ym1 = y[(long)(n - 1) * ys], ym2 = y[(long)(n - 2) * ys],
               ym3 = y[(long)(n - 3) * ys];
  double const c1 = y1 - (y0 - 2.0 * y1 + y2) / 6.0;
  double const cm2 = ym2 - (ym3 - 2.0 * ym2 + ym1) / 6.0;
  // forward sweep over i = 2 .. n-3 (pivot beta_{i-1}); the rows read y_i
  // before C(i) overwrites it when y aliases c
  for (int i = 2; i <= n - 3; ++i) {
    double r = 6.0 * y[(long)i * ys];
    if (i == 2) r -= c1;
    if (i == n - 3) r -= cm2;
    if (i > 2) r -= C(i - 1) / spline_beta(i - 2);
    C(i) = r;
  }
  C(n - 3) = C(n - 3) / spline_beta(n - 4);
  for (int i = n - 4; i >= 2; --i) C(i) = (C(i) - C(i + 1)) / spline_beta(i - 1);
  C(1) = c1;
  C(n - 2) = cm2;
  C(0) = 6.0 * y1 - 4.0 * c1 - C(2);
  C(-1) = 6.0 * y0 - 4.0 * C(0) - c1;
  C(n - 1) = 6.0 * ym2 - 4.0 * cm2 - C(n - 3);
  C(n) = 6.0 * ym1 - 4.0 * C(n - 1) - cm2;
  C(-2) = 4.0 * C(-1) - 6.0 * C(0) + 4.0 * C(1) - C(2);
  C(n + 1) = 4.0 * C(n) - 6.0 * C(n - 1) + 4.0 * C(n - 2) - C(n - 3);
}

// The smoothing identity's remainder is below 2e-7 of the integral in every
// case measured (the 3-D count over exactly the table's z range, with its
// weight at the table's edge: 1.2e-7); the cubature and VEGAS add this bound,
// relative to the integral, to the error they report for the synthetic kinds
// when these read the smoothed HMF (and stop only when the sum meets the
// tolerance).
constexpr double kSmoothRel = 1e-6;
__host__ __device__ inline double smooth_bound(int kind, bool smooth)
{
  return smooth && kind <= Y3G_DSIGMA6D ? kSmoothRel : 0.0;
}

// 120 x the same weights as polynomials in f, lowest power first.
__device__ constexpr double kB5Poly[6][6] = {{1, -5, 10, -10, 5, -1},  {26, -50, 20, 20, -20, 5},
                                             {66, 0, -60, 0, 30, -10}, {26, 50, 20, -20, -20, 10},
                                             {1, 5, 10, 10, 5, -5},    {0, 0, 0, 0, 0, 1}};

// The centred quintic B-spline at f + 2, f + 1, f, f - 1, f - 2, f - 3: the
// weights of coefficients k = i - 2 .. i + 3 at fraction f of cell i.
template <typename T> Y3G_HD void b5_weights(T f, T* w)
{
  T const g = T(1) - f, f2 = f * f, f3 = f2 * f, f4 = f3 * f, f5 = f4 * f;
  T const s = T(1.0 / 120.0);
  w[0] = (g * g) * (g * g) * g * s;
  w[1] = (T(26) - T(50) * f + T(20) * f2 + T(20) * f3 - T(20) * f4 + T(5) * f5) * s;
  w[2] = (T(66) - T(60) * f2 + T(30) * f4 - T(10) * f5) * s;
  w[3] = (T(26) + T(50) * f + T(20) * f2 - T(20) * f3 - T(20) * f4 + T(10) * f5) * s;
  w[4] = (T(1) + T(5) * f + T(10) * f2 + T(10) * f3 + T(5) * f4 - T(5) * f5) * s;
  w[5] = f5 * s;
}

// The hat-smoothed HMF at (lnM, zt): cells and fractions clamped as the
// bilinear lookup clamps, times s (lnM log10 e - 13.81) + q.
template <typename T>
Y3G_HD T hmf_smooth(const PointDev& P, const T* c, double lnM, double zt)
{
  double u = (lnM - P.lnm0) * P.inv_dlnm;
  double v = (zt - P.z0) * P.inv_dz;
  u = u < 0.0 ? 0.0 : (u > double(P.nm - 1) ? double(P.nm - 1) : u);
  v = v < 0.0 ? 0.0 : (v > double(P.nz - 1) ? double(P.nz - 1) : v);
  int i = int(u), j = int(v);
  if (i > P.nm - 2) i = P.nm - 2;
  if (j > P.nz - 2) j = P.nz - 2;
  T wx[6], wz[6];
  b5_weights<T>(T(u - double(i)), wx);
  b5_weights<T>(T(v - double(j)), wz);
  int const ld = P.nm + 4;
  const T* c0 = c + j * ld + i;
  T h = T(0);
  for (int l = 0; l < 6; ++l) {
    const T* row = c0 + l * ld;
    T r = T(0);
    for (int k = 0; k < 6; ++k) r += wx[k] * row[k];
    h += wz[l] * r;
  }
  return h * T(P.p.hmf_s * (lnM * 0.4342944819 - 13.8124426028) + P.p.hmf_q);
}

// DV_DO_DZ_t::operator() (dv_do_dz_t.cuh:19-27).
template <typename T>
Y3G_HD T dv_do_dz(const PointDev& P, const T* da, double zt)
{
  T const d = lerp_clamp<T>(da, P.nda, zt, P.inv_dzda);
  T const zp = T(1) + T(zt);
  T const om = T(P.p.omega_m);
  T const ez = M<T>::sqrt(om * zp * zp * zp + (T(1) - om));
  T const h = T(P.p.h);
  return M<T>::div(T(kDvConst) * zp * zp * d * h * d * h, ez);
}

// NFW: F and h(x) = (1 - F) / (x^2 - 1), series in t = x^2 - 1 near x = 1.
template <typename T> Y3G_HD void nfw_F_h(T x, T& F, T& h)
{
  T const t = x * x - T(1);
  if (M<T>::fabs(t) < T(0.1)) {
    // h = sum_k (-t)^(k-1) / (2k+1), k >= 1
    T hs = T(0), tp = T(1);
    constexpr int nt = sizeof(T) == 4 ? 9 : 18;
#pragma unroll
    for (int k = 1; k <= nt; ++k) {
      T const term = tp / T(2 * k + 1);
      hs += (k & 1) ? term : -term;
      tp *= t;
    }
    h = hs;
    F = T(1) - t * hs;
    return;
  }
  if (t < T(0)) {
    T const s = M<T>::sqrt(-t);
    F = M<T>::div(M<T>::log(M<T>::div(T(1) + s, x)), s);
  } else {
    T const s = M<T>::sqrt(t);
    F = M<T>::div(M<T>::atan(s), s);
  }
  h = M<T>::div(T(1) - F, t);
}

#include "y3g_nfw_table.inc"

// FP32 form of nfw_g below as a function of x^2 by the knot table of
// y3g_nfw_table.inc (fit_nfw_table.py): cubic Hermite interpolation of
// v = log2(g (1 + x^2)) in s = log2(x^2), clamped to the table's ends
// (x from 1e-6 to 6.5e4); relative error below 7e-7, no branches.
__device__ __forceinline__ float nfw_g_tab(float x2, const NfwKnot* tab)
{
  float u = (__builtin_amdgcn_logf(x2) - kNfwS0) * kNfwInvStep;
  u = fminf(fmaxf(u, 0.0f), float(kNfwN - 1) - 1e-3f);
  int const i = int(u);
  float const t = u - float(i), t2 = t * t, t3 = t2 * t;
  NfwKnot const a = tab[i], b = tab[i + 1];
  float const v = (2.0f * t3 - 3.0f * t2 + 1.0f) * a.v + (t3 - 2.0f * t2 + t) * a.d +
                  (3.0f * t2 - 2.0f * t3) * b.v + (t3 - t2) * b.d;
  return __builtin_amdgcn_exp2f(v) * __builtin_amdgcn_rcpf(1.0f + x2);
}

// Delta-Sigma / (rho_s r_s) of NFW (Wright & Brainerd 2000).
template <typename T> Y3G_HD T nfw_g(T x)
{
  if constexpr (sizeof(T) == 4) return nfw_g_tab(x * x, kNfwTab);
  T F, h, m;
  nfw_F_h<T>(x, F, h);
  if (x < T(0.9)) {
    T const s = M<T>::sqrt(T(1) - x * x);
    T const q = M<T>::div(x * x, T(2) * (T(1) + s));
    T const phi = M<T>::div(-M<T>::log1p(-q), q);
    m = M<T>::div(M<T>::div(M<T>::log(M<T>::div(T(2), x)), T(1) + s) -
                      M<T>::div(T(0.5) * phi, T(1) + s),
                  s);
  } else {
    m = M<T>::div(M<T>::log(T(0.5) * x) + F, x * x);
  }
  return T(4) * m - T(2) * h;
}

// Scale radius, amplitude 1e-12 delta_c rho r_s of an NFW halo with
// Delta = 200 rho and c = conc (> 0) or c(M) = 5 (M/1e14)^-0.1
// (nfw_dsigma_mis.cuh:104-123 normalisation).
template <typename T> struct NfwHalo {
  T rs, amp;
};
template <typename T>
Y3G_HD NfwHalo<T> nfw_halo(double lnM, T rho, T conc)
{
  T const dl = T(lnM - kLnPivotHmf);
  T const c = conc > T(0) ? conc : T(5) * M<T>::exp(T(-0.1) * dl);
  // r200 = cbrt(3 M / (800 pi rho)), M = 1e14 e^dl, through logarithms
  T const r200 = M<T>::exp((dl + T(kLn3e14Over800Pi) - M<T>::log(rho)) * T(1.0 / 3.0));
  T const rs = M<T>::div(r200, c);
  T const dc = M<T>::div(T(200.0 / 3.0) * c * c * c,
                         M<T>::log(T(1) + c) - M<T>::div(c, T(1) + c));
  NfwHalo<T> out;
  out.rs = rs;
  out.amp = T(1e-12) * dc * rho * rs;
  return out;
}

template <typename T> Y3G_HD T nfw_dsigma(T R, const NfwHalo<T>& H)
{
  return H.amp * nfw_g<T>(M<T>::div(R, H.rs));
}

// Tangential Delta-Sigma about a centre offset by rmis at angle theta:
// R' (sigma_miscent_y1_mort_cuda_scalarintegrand.cu:176-178), cos(2 beta).
// FP32 form from 1 / r_s: R'^2 and x^2 = R'^2 / r_s^2 need no square root.
__device__ __forceinline__ float dsigma_offset_f(float R, float rmis, float theta, float irs,
                                                 float amp, const NfwKnot* tab)
{
  float const hc = __cosf(0.5f * theta);
  float const k = hc * hc;
  float const dr = R - rmis;
  float const rp2 = dr * dr + 4.0f * R * rmis * k;
  if (!(rp2 > 1e-20f)) return 0.0f;
  float const a = dr + 2.0f * rmis * k;
  float const c2b = 2.0f * a * a * __builtin_amdgcn_rcpf(rp2) - 1.0f;
  return amp * nfw_g_tab(rp2 * irs * irs, tab) * c2b;
}

template <typename T>
Y3G_HD T dsigma_offset(T R, T rmis, T theta, const NfwHalo<T>& H)
{
  if constexpr (sizeof(T) == 4)
    return dsigma_offset_f(R, rmis, theta, __builtin_amdgcn_rcpf(H.rs), H.amp, kNfwTab);
  // with k = cos^2(theta/2): R'^2 = (R - rmis)^2 + 4 R rmis k and
  // R + rmis cos(theta) = (R - rmis) + 2 rmis k, free of the cancellation
  // of the textbook form near R = rmis, theta = pi
  T const hc = M<T>::cos(T(0.5) * theta);
  T const k = hc * hc;
  T const dr = R - rmis;
  T const rp2 = dr * dr + T(4) * R * rmis * k;
  if (!(rp2 > T(1e-20))) return T(0);
  T const a = dr + T(2) * rmis * k;
  T const c2b = M<T>::div(T(2) * a * a, rp2) - T(1);
  return nfw_dsigma<T>(M<T>::sqrt(rp2), H) * c2b;
}

// ------------------------------------------------------- integrands
// Mapped integrand: y are the integration variables of the box.  The MOR
// variable is integrated in its standardised form
//   t = (ln lt - ln lbar(lnM, zt)) / sigma_ln,  MOR(lt) dlt = phi(t) dt,
// with t running over its window [t(lt_lo), t(lt_hi)] clipped smoothly to [-7, 7]
// through v in [0, 1]; for Delta-Sigma the offset radius is integrated in
// its Rayleigh CDF u (P(rmis) drmis = du) and the angle on [0, pi] (the
// integrand is even about pi).
// Axis positions of the integration variables per kind, at compile time
// (node coordinates then stay in registers).
template <int K> struct KindAxes;
template <> struct KindAxes<Y3G_COUNT3D> {
  static constexpr int it = 0, iz = 1, im = 2, nd = 3;
};
template <> struct KindAxes<Y3G_COUNT5D> {
  static constexpr int it = 2, iz = 3, im = 4, nd = 5;
};
template <> struct KindAxes<Y3G_DSIGMA6D> {
  static constexpr int it = 1, iz = 2, im = 3, nd = 6;
};
// the miscentred real-data kinds integrate lambda_obs in closed form
// (lo_lc_bin): (lc, zt, lnM, rmis, theta) and (lc, zt, lnM, rmis)
template <> struct KindAxes<Y3G_DSIGMA6D_DES> {
  static constexpr int it = 0, iz = 1, im = 2, nd = 5;
};
template <> struct KindAxes<Y3G_DSIGMA3D_DES> {
  static constexpr int it = 0, iz = 1, im = 2, nd = 3;
};
template <> struct KindAxes<Y3G_COUNT3D_DES> {
  static constexpr int it = 0, iz = 1, im = 2, nd = 3;
};
template <> struct KindAxes<Y3G_COUNT5D_DES> {
  static constexpr int it = 0, iz = 1, im = 2, nd = 4;
};
template <> struct KindAxes<Y3G_COUNT4D_CARD> {
  static constexpr int it = 1, iz = 2, im = 3, nd = 4;
};
template <> struct KindAxes<Y3G_POP_P1> {
  static constexpr int it = 0, iz = 1, im = 2, nd = 3;
};
template <> struct KindAxes<Y3G_POP_I1> {
  static constexpr int it = 0, iz = 1, im = 2, nd = 3;
};
template <> struct KindAxes<Y3G_POP_I2> {
  static constexpr int it = 0, iz = 1, im = 2, nd = 3;
};
template <> struct KindAxes<Y3G_POP_J> {
  static constexpr int it = 0, iz = 1, im = 2, nd = 3;
};
// The FP64-only kinds that read per-point data (y3g_des_upload): the
// real-data DES integrands and the P[X] operator integrands.
constexpr bool is_des_kind(int k) { return k >= Y3G_DSIGMA6D_DES && k <= Y3G_POP_J; }
constexpr bool is_pop_kind(int k) { return k >= Y3G_POP_P1 && k <= Y3G_POP_J; }
// DES Y3 richness bin centres of the P[X] operator (p_operator_gpu_t.cuh:55-60)
__host__ __device__ inline double pop_lob_center(int bin)
{
  return bin == 1 ? 37.5 : bin == 2 ? 52.5 : bin == 3 ? 130.0 : 25.0;
}

// The MOR variable's window: [a0, b0] = [t(lt_lo), t(lt_hi)] clipped to
// [-T, T] by a smoothed clip, tlo = a0 + softplus(k (-T - a0)) / k and
// thi = b0 - softplus(k (b0 - T)) / k (k = kTSoft).  tlo lies within
// ln 2 / k of max(a0, -T) and never below it (likewise thi), so the window
// drops < 1e-11 of the Gaussian mass, while a hard max / min would give the
// integrand a kink along the curves a0(lnM, zt) = -T and b0(lnM, zt) = T.
template <typename T> Y3G_HD void t_window(T a0, T b0, T& tlo, T& thi)
{
  T const ua = T(kTSoft) * (T(-kTMax) - a0), ub = T(kTSoft) * (b0 - T(kTMax));
  T const sa = ua > T(30) ? ua : M<T>::log1p(M<T>::exp(ua));
  T const sb = ub > T(30) ? ub : M<T>::log1p(M<T>::exp(ub));
  tlo = a0 + sa * T(1.0 / kTSoft);
  thi = b0 - sb * T(1.0 / kTSoft);
}

// z_obs of the 5-D count through its standardised photo-z variable
// w = (zo - zt) / sz, sz = sigz0 (1 + zt), over the window of [zo_lo, zo_hi]
// clipped smoothly to |w| <= 7 (t_window), as v in [0, 1]:
// P(zo | zt) dzo = phi(w) (w_hi - w_lo) dv.  The change of variables is exact
// for any window inside the bin (the clip drops < 1e-11 of the Gaussian), and
// it straightens the zo = zt ridge of width sz (0.01 - 0.017, a tenth of the
// bin) into a fixed bump in v, which axis-aligned regions resolve.
template <typename T> Y3G_HD T zo_std(T v, T zt, T zlo, T zhi, T sigz0)
{
  T const isz = M<T>::div(T(1), sigz0 * (T(1) + zt));
  T wlo, whi;
  t_window<T>((zlo - zt) * isz, (zhi - zt) * isz, wlo, whi);
  if (!(whi > wlo)) return T(0);
  T const w = wlo + (whi - wlo) * v;
  return M<T>::exp(T(-0.5) * w * w) * T(0.3989422804014327) * (whi - wlo);
}

// Integrand of kind K (axis positions known at compile time).
template <typename T, int K>
Y3G_HD T integrand_k(const y3g_integral& ig, const PointDev& P, Tabs<T> tb,
                     const double* y)
{
  constexpr int it = KindAxes<K>::it, iz = KindAxes<K>::iz, im = KindAxes<K>::im;
  double const zt = y[iz], lnM = y[im];
  T const sln = T(P.p.mor_sigma_ln);
  T const zp = T(1) + T(zt);
  T const lnlbar = T(P.p.mor_lnl0) + T(P.p.mor_slope) * T(lnM - kLnPivotMor) +
                   T(P.p.mor_zslope) * M<T>::log(zp * T(1.0 / 1.35));
  T const isl = M<T>::div(T(1), sln);
  T tlo, thi;
  t_window<T>((T(ig.lnlt_lo) - lnlbar) * isl, (T(ig.lnlt_hi) - lnlbar) * isl, tlo, thi);
  if (!(thi > tlo)) return T(0);
  T const tt = tlo + (thi - tlo) * T(y[it]);
  T const lt = M<T>::exp(lnlbar + sln * tt);
  T const phi = M<T>::exp(T(-0.5) * tt * tt) * T(0.3989422804014327) * (thi - tlo);
  T const hm = tb.hmfc ? hmf_smooth<T>(P, tb.hmfc, lnM, zt) : hmf<T>(P, tb.hmf, lnM, zt);
  T const common = hm * phi * dv_do_dz<T>(P, tb.da, zt);
  T const fprj = T(P.p.f_prj), pscale = T(P.p.prj_scale);
  if constexpr (K == Y3G_COUNT3D) {
    return common * lo_lt_bin<T>(T(ig.lo_lo), T(ig.lo_hi), lt, T(0), pscale);
  } else if constexpr (K == Y3G_COUNT5D) {
    T const lo = T(y[0]);
    return common * lo_lt_pdf<T>(lo, lt, fprj, pscale) *
           zo_std<T>(T(y[1]), T(zt), T(ig.zo_lo), T(ig.zo_hi), T(P.p.sigz0));
  } else {
    T const lo = T(y[0]);
    T const s = T(ig.sigma_mis);
    // -2 ln(1 - u): log1p for small u, ln of the FP64 difference near 1
    double const u = y[4];
    T const l1mu = u < 0.5 ? M<T>::log1p(T(-u)) : M<T>::log(T(1.0 - u));
    T const rmis = s * M<T>::sqrt(T(-2) * l1mu);
    T const th = T(y[5]);
    T const rho = T(P.rho_m);
    NfwHalo<T> const H = nfw_halo<T>(lnM, rho, T(0));
    T const R = T(ig.radius);
    T const fmis = T(P.p.f_mis);
    T const prof = (T(1) - fmis) * nfw_dsigma<T>(R, H) +
                   fmis * dsigma_offset<T>(R, rmis, th, H);
    T const sci = lerp_clamp<T>(tb.sci, P.nsci, zt, P.inv_dzsci);
    // (1/2pi) x 2 (theta symmetry) x rmis_norm (truncated Rayleigh)
    return common * lo_lt_pdf<T>(lo, lt, fprj, pscale) *
           zo_bin<T>(T(ig.zo_lo), T(ig.zo_hi), T(zt), T(P.p.sigz0)) * sci *
           T(ig.rmis_norm * 0.31830988618379067) * prof;
  }
}

template <typename T>
Y3G_HD T integrand(const y3g_integral& ig, const PointDev& P, Tabs<T> tb,
                   const double* y)
{
  switch (ig.kind) {
    case Y3G_COUNT3D: return integrand_k<T, Y3G_COUNT3D>(ig, P, tb, y);
    case Y3G_COUNT5D: return integrand_k<T, Y3G_COUNT5D>(ig, P, tb, y);
    default: return integrand_k<T, Y3G_DSIGMA6D>(ig, P, tb, y);
  }
}

// The same integrand split in two: the factors that depend on (zt, lnM)
// only, and the rest.  A Genz-Malik region's nodes take only 21 distinct
// (zt, lnM) pairs, so the cubature evaluates pair_factors once per pair
// and integrand_pair once per node.
template <typename T> struct PairF {
  T lnlbar, tlo, dt;  // MOR centre, the clipped t window [tlo, tlo + dt]
  T zfac;             // hmf x dV/dz x dt / sqrt(2 pi)
                      //   [x zo bin x Sigma_crit^-1 x angular norm]
  T irs, amp;         // NFW halo (Delta-Sigma): 1 / r_s, amplitude
  T cds;              // (1 - f_mis) x centred Delta-Sigma(R)
};

template <typename T>
Y3G_HD void kind_axes(int kind, int& it, int& iz, int& im)
{
  switch (kind) {
    case Y3G_COUNT3D: it = 0, iz = 1, im = 2; break;
    case Y3G_COUNT5D: it = 2, iz = 3, im = 4; break;
    default: it = 1, iz = 2, im = 3; break;
  }
}

// The pair factors themselves split once more: what depends on zt alone
// (the MOR redshift term, dV/dz [x zo bin x Sigma_crit^-1], the HMF table
// row), what depends on lnM alone (the MOR mass term, the HMF table column
// and amplitude factor, the NFW halo and the centred profile), and their
// combination (the bilinear HMF value, the t window).  A region's 21 pairs
// take 7 distinct zt and 7 distinct lnM values.
template <typename T> struct ZFac {
  T lz, zz, sz;  // MOR z term, dV/dz [x ...], HMF row fraction
  int jz;        // HMF row
};
template <typename T> struct MFac {
  T lm, hamp, tm;    // MOR mass term, HMF amplitude factor, column fraction
  T irs, amp, cds;   // NFW halo, centred profile (Delta-Sigma)
  int im;            // HMF column
};

template <typename T>
Y3G_HD ZFac<T> z_factors(const y3g_integral& ig, const PointDev& P, Tabs<T> tb, double zt)
{
  ZFac<T> z;
  T const zp = T(1) + T(zt);
  z.lz = T(P.p.mor_zslope) * M<T>::log(zp * T(1.0 / 1.35));
  double v = (zt - P.z0) * P.inv_dz;
  v = v < 0.0 ? 0.0 : (v > double(P.nz - 1) ? double(P.nz - 1) : v);
  int j = int(v);
  if (j > P.nz - 2) j = P.nz - 2;
  z.jz = j;
  z.sz = T(v - double(j));
  T zz = dv_do_dz<T>(P, tb.da, zt);
  if (ig.kind == Y3G_DSIGMA6D) {
    T const sci = lerp_clamp<T>(tb.sci, P.nsci, zt, P.inv_dzsci);
    // (1/2pi) x 2 (theta symmetry) x rmis_norm (truncated Rayleigh)
    zz *= zo_bin<T>(T(ig.zo_lo), T(ig.zo_hi), T(zt), T(P.p.sigz0)) * sci *
          T(ig.rmis_norm * 0.31830988618379067);
  }
  z.zz = zz;
  return z;
}

template <typename T>
Y3G_HD MFac<T> m_factors(const y3g_integral& ig, const PointDev& P, double lnM,
                         const NfwKnot* nfw)
{
  MFac<T> m;
  m.lm = T(P.p.mor_lnl0) + T(P.p.mor_slope) * T(lnM - kLnPivotMor);
  double u = (lnM - P.lnm0) * P.inv_dlnm;
  u = u < 0.0 ? 0.0 : (u > double(P.nm - 1) ? double(P.nm - 1) : u);
  int i = int(u);
  if (i > P.nm - 2) i = P.nm - 2;
  m.im = i;
  m.tm = T(u - double(i));
  m.hamp = T(P.p.hmf_s * (lnM * 0.4342944819 - 13.8124426028) + P.p.hmf_q);
  m.irs = m.amp = m.cds = T(0);
  if (ig.kind == Y3G_DSIGMA6D) {
    NfwHalo<T> const H = nfw_halo<T>(lnM, T(P.rho_m), T(0));
    m.irs = M<T>::div(T(1), H.rs);
    m.amp = H.amp;
    if constexpr (sizeof(T) == 4) {
      float const x = float(ig.radius) * m.irs;
      m.cds = (1.0f - float(P.p.f_mis)) * (H.amp * nfw_g_tab(x * x, nfw));
    } else {
      m.cds = (T(1) - T(P.p.f_mis)) * nfw_dsigma<T>(T(ig.radius), H);
    }
  }
  return m;
}

// SM: the smoothed HMF (tb.hmfc) at compile time; the FP64 direct integrand
// and this function's other callers pass false and read the bilinear table.
template <typename T, bool SM = false>
Y3G_HD PairF<T> pair_combine(const y3g_integral& ig, const PointDev& P, Tabs<T> tb,
                             const ZFac<T>& z, const MFac<T>& m)
{
  PairF<T> f;
  f.lnlbar = m.lm + z.lz;
  T const isl = M<T>::div(T(1), T(P.p.mor_sigma_ln));
  T tlo, thi;
  t_window<T>((T(ig.lnlt_lo) - f.lnlbar) * isl, (T(ig.lnlt_hi) - f.lnlbar) * isl, tlo, thi);
  f.tlo = tlo;
  f.dt = thi - tlo;
  // HMF: clamped bilinear lookup (hmf_t.cuh:54-60), as hmf() below; the
  // table is in global memory for every caller
  typedef const __attribute__((address_space(1))) T* GPtr;
  T hm;
  if constexpr (SM) {
    // hat-smoothed HMF (hmf_smooth): the 6 x 6 coefficients around the cell
    // row by row, three rows' loads in flight at a time (unrolling all six
    // spills more and ran slower), the row's z weight by Horner from kB5Poly
    T wx[6];
    b5_weights<T>(m.tm, wx);
    T const fz = z.sz;
    int const ld = P.nm + 4;
    GPtr row = (GPtr)(tb.hmfc + z.jz * ld + m.im);
    T h = T(0);
#pragma unroll 3
    for (int l = 0; l < 6; ++l, row += ld) {
      T r = wx[0] * row[0];
#pragma unroll
      for (int k = 1; k < 6; ++k) r += wx[k] * row[k];
      const double* a = kB5Poly[l];
      T w = T(a[5]);
#pragma unroll
      for (int q = 4; q >= 0; --q) w = w * fz + T(a[q]);
      h += w * r;
    }
    hm = h * (m.hamp * T(1.0 / 120.0));
  } else {
    GPtr r0 = (GPtr)(tb.hmf + z.jz * P.nm + m.im);
    GPtr r1 = r0 + P.nm;
    T const a = r0[0] + m.tm * (r0[1] - r0[0]);
    T const b = r1[0] + m.tm * (r1[1] - r1[0]);
    hm = (a + z.sz * (b - a)) * m.hamp;
  }
  f.zfac = hm * z.zz * (T(0.3989422804014327) * f.dt);
  f.irs = m.irs;
  f.amp = m.amp;
  f.cds = m.cds;
  return f;
}

template <typename T>
Y3G_HD PairF<T> pair_factors(const y3g_integral& ig, const PointDev& P, Tabs<T> tb,
                             double zt, double lnM)
{
  return pair_combine<T>(ig, P, tb, z_factors<T>(ig, P, tb, zt),
                         m_factors<T>(ig, P, lnM, kNfwTab));
}

template <typename T, int K>
Y3G_HD T integrand_pair_k(const y3g_integral& ig, const PointDev& P, const PairF<T>& f,
                          const double* y)
{
  if (!(f.dt > T(0))) return T(0);
  constexpr int it = KindAxes<K>::it, iz = KindAxes<K>::iz;
  T const sln = T(P.p.mor_sigma_ln);
  T const tt = f.tlo + f.dt * T(y[it]);
  T const lt = M<T>::exp(f.lnlbar + sln * tt);
  T const common = f.zfac * M<T>::exp(T(-0.5) * tt * tt);
  T const fprj = T(P.p.f_prj), pscale = T(P.p.prj_scale);
  if constexpr (K == Y3G_COUNT3D) {
    return common * lo_lt_bin<T>(T(ig.lo_lo), T(ig.lo_hi), lt, T(0), pscale);
  } else if constexpr (K == Y3G_COUNT5D) {
    T const lo = T(y[0]);
    return common * lo_lt_pdf<T>(lo, lt, fprj, pscale) *
           zo_std<T>(T(y[1]), T(y[iz]), T(ig.zo_lo), T(ig.zo_hi), T(P.p.sigz0));
  } else {
    T const lo = T(y[0]);
    T const s = T(ig.sigma_mis);
    double const u = y[4];
    T const l1mu = u < 0.5 ? M<T>::log1p(T(-u)) : M<T>::log(T(1.0 - u));
    T const rmis = s * M<T>::sqrt(T(-2) * l1mu);
    NfwHalo<T> H;
    H.rs = M<T>::div(T(1), f.irs);
    H.amp = f.amp;
    T const prof = f.cds + T(P.p.f_mis) * dsigma_offset<T>(T(ig.radius), rmis, T(y[5]), H);
    return common * lo_lt_pdf<T>(lo, lt, fprj, pscale) * prof;
  }
}

// Per-integral constants of the FP32 node integrand, converted once per
// chunk of regions and kept in registers across its nodes.
struct NodeConstF {
  float sln, fprj, pscale, ips, sigz0, smis, fmis, radius, lo_lo, lo_hi, zo_lo, zo_hi;
  const NfwKnot* nfw;  // the NFW knot table (LDS copy in the cubature)
};
Y3G_HD NodeConstF node_const_f(const y3g_integral& ig, const PointDev& P, const NfwKnot* nfw)
{
  NodeConstF k;
  k.sln = float(P.p.mor_sigma_ln);
  k.fprj = float(P.p.f_prj);
  k.pscale = float(P.p.prj_scale);
  k.ips = float(1.0 / P.p.prj_scale);
  k.sigz0 = float(P.p.sigz0);
  k.smis = float(ig.sigma_mis);
  k.fmis = float(P.p.f_mis);
  k.radius = float(ig.radius);
  k.lo_lo = float(ig.lo_lo);
  k.lo_hi = float(ig.lo_hi);
  k.zo_lo = float(ig.zo_lo);
  k.zo_hi = float(ig.zo_hi);
  k.nfw = nfw;
  return k;
}

// FP32 node form of integrand_pair_k: the node's coordinates in float (yf)
// and, for Delta-Sigma, 1 - u in float (w) next to u (yf[4]); each integrand
// variable then costs one FMA from its region's centre and half-width.
// The FP32 cubature's per-pair factors: those the nodes read.
struct PairC {
  float lnlbar, tlo, dt, zfac;
};
Y3G_HD PairC pair_c(const PairF<float>& f)
{
  return PairC{f.lnlbar, f.tlo, f.dt, f.zfac};
}

// Miscentred profile of Delta-Sigma, f_mis x offset + (1 - f_mis) x centred
// (cds), at the Rayleigh CDF value u (w = 1 - u, accurate near u = 1) and the
// angle theta, for a halo of 1 / r_s = irs and amplitude amp.  It depends on
// (lnM, u, theta) only: a region's 149 nodes take 33 distinct triples.
Y3G_HD float profile_f(const NodeConstF& kc, float irs, float amp, float cds, float u, float w,
                       float theta)
{
  using T = float;
  // ln(1 - u): of w for u >= 1/2, else log1p(-u) in Goldberg's form; one
  // logarithm either way
  bool const low = u < T(0.5);
  T const arg = low ? T(1) - u : w;
  T const L = M<T>::log(arg);
  T const l1mu = !low ? L : (arg == T(1) ? -u : L * (-u) * __builtin_amdgcn_rcpf(arg - T(1)));
  T const rmis = kc.smis * M<T>::sqrt(T(-2) * l1mu);
  return cds + kc.fmis * dsigma_offset_f(kc.radius, rmis, theta, irs, amp, kc.nfw);
}

// The factor of the 5-D count and Delta-Sigma integrands on (lo, t, zt,
// lnM): the pair's zfac exp(-t^2 / 2) times the line-of-sight pdf at lo.  A
// region's 93 / 149 nodes take 57 distinct values of these four coordinates.
Y3G_HD float quad_f(const NodeConstF& kc, const PairC& f, float lo, float yt)
{
  using T = float;
  if (!(f.dt > T(0))) return T(0);
  T const tt = f.tlo + f.dt * yt;
  T const lt = M<T>::exp(f.lnlbar + kc.sln * tt);
  T const common = f.zfac * M<T>::exp(T(-0.5) * tt * tt);
  return common * lo_lt_pdf_f(lo, lt, kc.fprj, kc.ips);
}

// FP32 node form of integrand_pair_k: the node's coordinates in float (yf;
// lo, t [, zo]), its pair's factors and, for Delta-Sigma, its profile.  The
// cubature evaluates the same products factor by factor (quad_f, profile_f,
// zo_pdf).
template <int K>
Y3G_HD float integrand_node_f(const NodeConstF& kc, const PairC& f, const float* yf, float prof)
{
  using T = float;
  constexpr int it = KindAxes<K>::it, iz = KindAxes<K>::iz;
  if constexpr (K == Y3G_COUNT3D) {
    if (!(f.dt > T(0))) return T(0);
    T const tt = f.tlo + f.dt * yf[it];
    T const lt = M<T>::exp(f.lnlbar + kc.sln * tt);
    T const common = f.zfac * M<T>::exp(T(-0.5) * tt * tt);
    return common * lo_lt_bin<T>(kc.lo_lo, kc.lo_hi, lt, T(0), kc.pscale);
  } else if constexpr (K == Y3G_COUNT5D) {
    return quad_f(kc, f, yf[0], yf[it]) * zo_std<T>(yf[1], yf[iz], kc.zo_lo, kc.zo_hi, kc.sigz0);
  } else {
    return quad_f(kc, f, yf[0], yf[it]) * prof;
  }
}

template <typename T>
Y3G_HD T integrand_pair(const y3g_integral& ig, const PointDev& P, const PairF<T>& f,
                        const double* y)
{
  switch (ig.kind) {
    case Y3G_COUNT3D: return integrand_pair_k<T, Y3G_COUNT3D>(ig, P, f, y);
    case Y3G_COUNT5D: return integrand_pair_k<T, Y3G_COUNT5D>(ig, P, f, y);
    default: return integrand_pair_k<T, Y3G_DSIGMA6D>(ig, P, f, y);
  }
}

// ------------------------------------------- fixed-GL counts with selection
// Effective survey area Omega(z) in rad^2 (omega_z_des.hh:15-33): three
// degree-5 polynomials, coefficients leading power first; the middle piece
// in z - 0.6.
Y3G_HD double omega_z_des(double z)
{
  double c[6];
  double x = z;
  if (z < 0.504) {
    double const a[6] = {0.0, 0.0, 0.0, -0.00262353, 0.01940118, 0.45133063};
    for (int i = 0; i < 6; ++i) c[i] = a[i];
  } else if (z < 0.7) {
    double const a[6] = {1.33647377e+4, 1.35291046e+3, -1.26204891e+2,
                         -2.83454918e+1, -2.26465905, 3.84958753e-1};
    for (int i = 0; i < 6; ++i) c[i] = a[i];
    x = z - 0.6;
  } else {
    double const a[6] = {0.0, 0.0, -1.88101967, 4.8071839, -4.11424324, 1.18196785};
    for (int i = 0; i < 6; ++i) c[i] = a[i];
  }
  double out = 0.0;
  for (int i = 0; i < 6; ++i) out = c[i] + x * out;
  return out;
}

// Largest i in [0, n - 2] with xs[i] <= x (xs increasing, x inside).
Y3G_HD int bracket(const double* xs, int n, double x)
{
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    int const mid = (lo + hi) >> 1;
    if (xs[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

// Selection function S(lnM, z) of one bin (sel_function_t.hh:28-35): zero
// outside the axes, else bilinear in the (n_z, n_lnm) plane, lnM fastest.
Y3G_HD double sel_function(const double* lnm, int nm, const double* z, int nz,
                           const double* plane, double lnM, double zt)
{
  if (lnM < lnm[0] || lnM > lnm[nm - 1] || zt < z[0] || zt > z[nz - 1]) return 0.0;
  int const i = bracket(lnm, nm, lnM), j = bracket(z, nz, zt);
  double const t = (lnM - lnm[i]) / (lnm[i + 1] - lnm[i]);
  double const u = (zt - z[j]) / (z[j + 1] - z[j]);
  const double* r0 = plane + j * nm + i;
  const double* r1 = r0 + nm;
  return (1.0 - t) * (1.0 - u) * r0[0] + t * (1.0 - u) * r0[1] + (1.0 - t) * u * r1[0] +
         t * u * r1[1];
}

// INT_LC_LT_DES_t2::operator() (int_lc_lt_des_t2.cuh:48-63): the DES
// richness transfer P(lambda_c | lambda_true, z_true) from one of four
// (lt, zt) tables, picked by the lambda_c bin [20, 30), [30, 45), [45, 60),
// and the fourth table otherwise (lambda_c < 20 included).  The tables are
// the reference's data files read in file order into Interp2D(xs = lt knots,
// ys = zt knots, zs), lt fastest: value(lt_i, zt_j) = tab[j n_lt + i];
// bilinear between knots.  Packed as
//   par = {n_lt, n_zt, lt[n_lt], zt[n_zt], tab0, tab1, tab2, tab3}.
// Outside the knots every table is clamped to its edge (the reference clamps
// the fourth; the first three go through the device Interp2D's operator() of
// the un-vendored gpuintegration library).
Y3G_HD double int_lc_lt_des(const double* par, double lc, double lt, double zt)
{
  int const nl = int(par[0]), nz = int(par[1]);
  const double* xl = par + 2;
  const double* xz = xl + nl;
  int const b = (lc >= 20.0 && lc < 30.0)   ? 0
                : (lc >= 30.0 && lc < 45.0) ? 1
                : (lc >= 45.0 && lc < 60.0) ? 2
                                            : 3;
  const double* tab = xz + nz + (size_t)b * nl * nz;
  double const x = lt < xl[0] ? xl[0] : (lt > xl[nl - 1] ? xl[nl - 1] : lt);
  double const y = zt < xz[0] ? xz[0] : (zt > xz[nz - 1] ? xz[nz - 1] : zt);
  int const i = bracket(xl, nl, x), j = bracket(xz, nz, y);
  double const t = (x - xl[i]) / (xl[i + 1] - xl[i]);
  double const u = (y - xz[j]) / (xz[j + 1] - xz[j]);
  const double* r0 = tab + j * nl + i;
  const double* r1 = r0 + nl;
  return (1.0 - t) * (1.0 - u) * r0[0] + t * (1.0 - u) * r0[1] + (1.0 - t) * u * r1[0] +
         t * u * r1[1];
}

// MOR_DES_t::operator() (mor_des_t.cuh:276-305): the DES skew-normal
// mass-observable relation P(lambda_true | lnM, zt).
//   lambda(M, z) = max(1, ((M - A) / (B - A))^alpha ((1 + z) / (1 + z_p))^eps)
// (a mass below A gives a NaN base and lambda = 1, as std::max(1.0, NaN)
// does); sigma and skewness are clamped bilinear lookups in the
// (sigma_intr, lambda) tables; with x = lt - lambda the value is
//   N(x; 0, sigma) erfc(-skw x / (sqrt(2) sigma)).
// q = {A, B, alpha, sigma_intr, epsilon, z_pivot} (cluster_abundance/mor_A,
// mor_B, mor_alpha, mor_sigma, mor_epsilon, z_mor_pivot).
#define Y3G_MOR_DES_TABLE(name, n) static __constant__ double name[n]
#include "y3g_mor_des_tables.inc"
#undef Y3G_MOR_DES_TABLE

// The MOR_DES knot axes and tables, wherever they are held (constant memory
// by default, a workgroup's LDS copy in the FP64 cubature).
struct MorDesTabs {
  const double* si;  // sigma_intr knots [14]
  const double* ls;  // lambda_sat knots [42]
  const double* sg;  // sigma table [14][42]
  const double* sk;  // skewness table [14][42]
};
constexpr int kMorDesWords = 14 + 42 + 2 * 14 * 42;
__device__ inline MorDesTabs mor_des_const_tabs()
{
  return {kMorDesSigintr, kMorDesLsat, kMorDesSigma, kMorDesSkew};
}
// A copy laid out [si | ls | sg | sk] in `buf`.
__device__ inline MorDesTabs mor_des_tabs_at(const double* buf)
{
  return {buf, buf + 14, buf + 56, buf + 56 + 588};
}

// Bilinear weights of (sigma_intr, lambda) in the knot grid, clamped: one
// search serves both tables.
struct MorDesCell {
  int i, j;
  double t, u;
};
__device__ inline MorDesCell mor_des_cell(MorDesTabs const& m, double si, double lam)
{
  double const x = si < m.si[0] ? m.si[0] : (si > m.si[13] ? m.si[13] : si);
  double const y = lam < m.ls[0] ? m.ls[0] : (lam > m.ls[41] ? m.ls[41] : lam);
  MorDesCell c;
  c.i = bracket(m.si, 14, x);
  c.j = bracket(m.ls, 42, y);
  c.t = (x - m.si[c.i]) / (m.si[c.i + 1] - m.si[c.i]);
  c.u = (y - m.ls[c.j]) / (m.ls[c.j + 1] - m.ls[c.j]);
  return c;
}
__device__ inline double mor_des_at(const double* tab, MorDesCell const& c)
{
  const double* r0 = tab + c.i * 42 + c.j;  // row = sigma_intr knot
  const double* r1 = r0 + 42;
  double const t = c.t, u = c.u;
  return (1.0 - t) * (1.0 - u) * r0[0] + t * (1.0 - u) * r1[0] + (1.0 - t) * u * r0[1] +
         t * u * r1[1];
}

// MOR_DES_LOG_t::operator() (mor_des_log_t.cuh:284-316): the same skew
// normal with M_min = 10^q0 and M_1 = 10^q1 M_min, lambda = 1 below M_min,
// and max(0, lambda) in place of max(1, lambda).
// q = {log10 M_min, log10 (M_1 / M_min), alpha, sigma_intr, epsilon, z_pivot}.
__device__ inline double mor_des_log(const double* q, double lt, double lnM, double zt,
                                     MorDesTabs const& m = mor_des_const_tabs())
{
  double const mmin = pow(10.0, q[0]);
  double const m1 = pow(10.0, q[1]) * mmin;
  double const v = pow((exp(lnM) - mmin) / (m1 - mmin), q[2]) *
                   pow((1.0 + zt) / (1.0 + q[5]), q[4]);
  double const v1 = exp(lnM) < mmin ? 1.0 : v;
  double const lam = 0.0 < v1 ? v1 : 0.0;  // std::max(0.0, v1)
  MorDesCell const cell = mor_des_cell(m, q[3], lam);
  double const sigma = mor_des_at(m.sg, cell);
  double const skw = mor_des_at(m.sk, cell);
  double const x = lt - lam;
  double const z = x / sigma;
  return exp(-z * z / 2.0) * 0.3989422804014327 / sigma *
         erfc(-skw * x / (1.4142135623730951 * sigma));
}

__device__ inline double mor_des(const double* q, double lt, double lnM, double zt,
                                 MorDesTabs const& m = mor_des_const_tabs())
{
  double const v = pow((exp(lnM) - q[0]) / (q[1] - q[0]), q[2]) *
                   pow((1.0 + zt) / (1.0 + q[5]), q[4]);
  double const lam = 1.0 < v ? v : 1.0;
  MorDesCell const cell = mor_des_cell(m, q[3], lam);
  double const sigma = mor_des_at(m.sg, cell);
  double const skw = mor_des_at(m.sk, cell);
  double const x = lt - lam;
  double const z = x / sigma;
  return exp(-z * z / 2.0) * 0.3989422804014327 / sigma *
         erfc(-skw * x / (1.4142135623730951 * sigma));
}

// EMG_DES_t (emg_des_t.cuh:59-206): the DES observed-richness distribution
// P(lob | ltr, z), a Gaussian plus an exponentially modified Gaussian
// (projection) tail.  Its eight parameter curves are linear interpolations
// in z, clamped at the ends (quad::Interp1D::clamp over plob_ltr_params/z),
// packed as
//   par = {n, z[n], a_mu[n], b_mu[n], a_sig[n], b_sig[n], a_tau[n],
//          b_tau[n], a_fprj[n], b_fprj[n]}.
__device__ inline double emg_des_curve(const double* zs, const double* v, int n, double z)
{
  double const x = z < zs[0] ? zs[0] : (z > zs[n - 1] ? zs[n - 1] : z);
  int const i = bracket(zs, n, x);
  return v[i] + (x - zs[i]) * (v[i + 1] - v[i]) / (zs[i + 1] - zs[i]);
}

// get_params (:95-117): mu = a_mu + b_mu l, sigma = b_sig l^a_sig,
// tau = b_tau / l^a_tau, fprj = b_fprj / (1 + e^-l)^a_fprj, l = max(ltr, 0.5),
// then sigma, tau >= 1e-8 and fprj in [0, 1].
__device__ inline void emg_des_params(const double* par, double ltr, double z, double& mu,
                                      double& sigma, double& tau, double& fprj)
{
  int const n = int(par[0]);
  const double* zs = par + 1;
  double c[8];
  for (int k = 0; k < 8; ++k) c[k] = emg_des_curve(zs, zs + n * (k + 1), n, z);
  double const l = fmax(ltr, 0.5);
  mu = c[0] + c[1] * l;
  sigma = c[3] * pow(l, c[2]);
  tau = c[5] / pow(l, c[4]);
  double const denom = pow(1.0 + exp(-l), c[6]);
  fprj = c[7] / fmax(denom, 1e-300);
  sigma = fmax(sigma, 1e-8);
  tau = fmax(tau, 1e-8);
  fprj = fmax(0.0, fmin(1.0, fprj));
}

// The reference's scaled complementary error function (:37-51): exp(x^2)
// erfc(x) below |x| = 4, a three-term asymptotic series above (reflected
// for x < 0).
__device__ inline double emg_des_erfcx(double x)
{
  double const ax = fabs(x);
  if (ax < 4.0) return exp(x * x) * erfc(x);
  double const x2 = ax * ax;
  double r = (1.0 / 1.7724538509055159) / ax * (1.0 - 0.5 / x2 + 0.75 / (x2 * x2));
  if (x < 0.0) r = 2.0 * exp(x2) - r;
  return r;
}

// cdf() (:128-156): Phi((x - mu)/sigma) - fprj [tail], the tail through
// the erfcx form, clamped to [0, 1]; +-inf give 1 / 0.
__device__ inline double emg_des_cdf(const double* par, double lob, double ltr, double z)
{
  if (!isfinite(lob)) return lob > 0.0 ? 1.0 : 0.0;
  double mu, sigma, tau, fprj;
  emg_des_params(par, ltr, z, mu, sigma, tau, fprj);
  double const zs = (lob - mu) / sigma;
  double const gcdf = 0.5 * (1.0 + erf(zs * 0.7071067811865475));
  double const u = (tau * sigma - zs) * 0.7071067811865475;
  double const base = 0.5 * emg_des_erfcx(fabs(u)) * exp(-0.5 * zs * zs);
  double tail;
  if (u < 0.0) {
    double A = -tau * (lob - mu) + 0.5 * tau * tau * sigma * sigma;
    A = fmax(-700.0, fmin(700.0, A));
    tail = exp(A) - base;
  } else {
    tail = base;
  }
  double const r = gcdf - fprj * tail;
  return fmax(0.0, fmin(1.0, r));
}

// operator() (:160-178): (1 - fprj) N(lob; mu, sigma) + fprj EMG(lob).
__device__ inline double emg_des_pdf(const double* par, double lob, double ltr, double z)
{
  double mu, sigma, tau, fprj;
  emg_des_params(par, ltr, z, mu, sigma, tau, fprj);
  double const zz = (lob - mu) / sigma;
  double const g = exp(-zz * zz / 2.0) * 0.3989422804014327 / sigma;
  double const ea = 0.5 * tau * (2.0 * mu + tau * sigma * sigma - 2.0 * lob);
  double const eb = (mu + tau * sigma * sigma - lob) / (1.4142135623730951 * sigma);
  double emg = 0.0;
  if (ea > -700.0 && ea < 700.0) emg = 0.5 * tau * exp(ea) * erfc(eb);
  return (1.0 - fprj) * g + fprj * emg;
}

// Clamped bilinear lookup in a table over (xs[nx], ys[ny]) stored x-fastest,
// value(x_i, y_j) = z[j nx + i]: the convention of the reference's Interp2D
// for a flattened table (interp_2d.cc:115-127; the DES tables' expected
// values confirm it for the device class).
__device__ inline double interp2d_clamp(const double* xs, int nx, const double* ys, int ny,
                                        const double* z, double x, double y)
{
  double const xc = x < xs[0] ? xs[0] : (x > xs[nx - 1] ? xs[nx - 1] : x);
  double const yc = y < ys[0] ? ys[0] : (y > ys[ny - 1] ? ys[ny - 1] : y);
  int const i = bracket(xs, nx, xc), j = bracket(ys, ny, yc);
  double const t = (xc - xs[i]) / (xs[i + 1] - xs[i]);
  double const u = (yc - ys[j]) / (ys[j + 1] - ys[j]);
  const double* r0 = z + j * nx + i;
  const double* r1 = r0 + nx;
  return (1.0 - t) * (1.0 - u) * r0[0] + t * (1.0 - u) * r0[1] + (1.0 - t) * u * r1[0] +
         t * u * r1[1];
}

// ------------------------------------------------ tabulated LC_LT_t
// LC_LT_t::operator()(lc, lt, zt) (lc_lt_t.cuh:36-68): Costanzi et al. 2018
// eq. 15 with its five parameters clamped bilinear in (lt, zt) over the
// reference's fitted tables (lc_lt_t.cc:10-214; y3g_lc_lt_tables.inc):
//   (1 - fmsk)(1 - fprj) G(lc; mu, sigma)
//   + 1/2 ((1 - fmsk) fprj tau + fmsk fprj / lt) e^{tau(2 mu + tau sigma^2 - 2 lc)/2}
//       erfc((mu + tau sigma^2 - lc) / (sqrt2 sigma))
//   + 1/2 fmsk / lt (erfc((lc - mu) / (sqrt2 sigma)) - erfc((lc + lt - mu) / (sqrt2 sigma)))
//   - 1/2 fmsk fprj / lt e^{-tau lt} e^{tau(...)} erfc((mu + tau sigma^2 - lc - lt) / (sqrt2 sigma))
#define Y3G_LC_LT_TABLE(name, n) static __constant__ double name[n]
#include "y3g_lc_lt_tables.inc"
#undef Y3G_LC_LT_TABLE

__device__ inline double lc_lt_tab(double lc, double lt, double zt)
{
  double const tau = interp2d_clamp(kLcLtLt, 22, kLcLtZt, 5, kLcLtTau, lt, zt);
  double const mu = interp2d_clamp(kLcLtLt, 22, kLcLtZt, 5, kLcLtMu, lt, zt);
  double const sigma = interp2d_clamp(kLcLtLt, 22, kLcLtZt, 5, kLcLtSigma, lt, zt);
  double const fmsk = interp2d_clamp(kLcLtLt, 22, kLcLtZt, 5, kLcLtFmsk, lt, zt);
  double const fprj = interp2d_clamp(kLcLtLt, 22, kLcLtZt, 5, kLcLtFprj, lt, zt);
  double const exptau = exp(tau * (2.0 * mu + tau * sigma * sigma - 2.0 * lc) / 2.0);
  double const r2s = 1.4142135623730951 * sigma;
  double const mts = mu + tau * sigma * sigma;
  double const z = (lc - mu) / sigma;
  double const gauss = exp(-0.5 * z * z) * 0.3989422804014327 / sigma;
  return (1.0 - fmsk) * (1.0 - fprj) * gauss +
         0.5 * ((1.0 - fmsk) * fprj * tau + fmsk * fprj / lt) * exptau * erfc((mts - lc) / r2s) +
         0.5 * fmsk / lt * (erfc((lc - mu) / r2s) - erfc((lc + lt - mu) / r2s)) -
         0.5 * fmsk * fprj / lt * (exp(-tau * lt) * exptau * erfc((mts - (lc + lt)) / r2s));
}

// ------------------------------------- tabulated miscentred NFW profile
// NFW_DSIGMA_MIS (nfw_dsigma_mis.cuh:51-131) reads T(ln x, ln xmis), the log
// of the gamma-kernel miscentred Delta-Sigma in units of 2 rho_s r_s, on
// 1000 x 250 log-uniform knots (x = r / r_s over [1e-3, 5e3], xmis = rmis /
// r_s over [1e-2, 5e2]; the reference's data/nfw_off_center/table_1000_
// 1e-03_5e+03_gamma_{logx,logxmis}.txt axes).  The table itself is not in the
// reference tree; it is built here from the closed-form profile:
//   T = ln(G / 2),  G(x, m) = int_0^60 u e^-u Ring(x, m u) du,
//   Ring(x, s) = (1/pi) int_0^pi g(X) cos 2 beta dphi,
//   X^2 = (x - s)^2 + 4 x s sin^2(phi/2),  cos beta = (x - s cos phi) / X,
// both integrals by 8-point Gauss-Legendre on graded panels (phi toward 0,
// where the offset centre comes closest; u toward u* = x / m from both
// sides, pieces at most 2 wide).
constexpr int kNfwMisNx = 1000, kNfwMisNm = 250;
constexpr int kNfwMisMaxPanels = 128;

__host__ __device__ inline double gl8_node(int k)
{
  constexpr double x[4] = {0.1834346424956498, 0.5255324099163290, 0.7966664774136267,
                           0.9602898564975363};
  return k < 4 ? -x[3 - k] : x[k - 4];
}
__host__ __device__ inline double gl8_weight(int k)
{
  constexpr double w[4] = {0.3626837833783620, 0.3137066458778873, 0.2223810344533745,
                           0.1012285362903763};
  return k < 4 ? w[3 - k] : w[k - 4];
}

// Panel levels of Ring(x, s): the first panel [0, pi 2^-L] is at most 1/16
// of the width phi_c = |x - s| / sqrt(x s) of the feature near phi = 0.
__host__ __device__ inline int nfw_ring_levels(double x, double s)
{
  double const pc = fabs(x - s) / sqrt(x * s);
  int L = 1;
  if (pc < 3.14159265358979323846) {
    double const q = log2(3.14159265358979323846 / fmax(pc, 1e-300));
    L = q > 36.0 ? 40 : (int)ceil(q) + 4;
  }
  return L > 40 ? 40 : L;
}

__host__ __device__ inline double nfw_ring(double x, double s)
{
  int const L = nfw_ring_levels(x, s);
  double acc = 0.0, a = 0.0;
  for (int j = 0; j <= L; ++j) {
    double const b = 3.14159265358979323846 * ldexp(1.0, j - L);
    double const c = 0.5 * (a + b), h = 0.5 * (b - a);
    double ps = 0.0;
    for (int k = 0; k < 8; ++k) {
      double const phi = c + h * gl8_node(k);
      double const sn = sin(0.5 * phi);
      double const d = x - s;
      double const X2 = d * d + 4.0 * x * s * sn * sn;
      double const cb = d + 2.0 * s * sn * sn;  // x - s cos(phi)
      ps += gl8_weight(k) * nfw_g<double>(sqrt(X2)) * (2.0 * cb * cb / X2 - 1.0);
    }
    acc += h * ps;
    a = b;
  }
  return acc / 3.14159265358979323846;
}

// The outer panels of G(x, m) on u in [0, 60] (at most kNfwMisMaxPanels).
__host__ __device__ inline int nfw_mis_panels(double x, double m, double* lo, double* hi)
{
  double const U = 60.0, us = x / m;
  double br[64];
  int nb = 0;
  br[nb++] = 0.0;
  if (us > 0.0 && us < U) {
    for (int k = 0; k <= 24; ++k) {
      double const p = us - us * ldexp(1.0, -k);
      if (p > br[nb - 1]) br[nb++] = p;
    }
    br[nb++] = us;
    for (int k = 24; k >= 0; --k) {
      double const p = us + (U - us) * ldexp(1.0, -k);
      if (p > br[nb - 1] && p <= U) br[nb++] = p;
    }
    if (br[nb - 1] < U) br[nb++] = U;
  } else {
    br[nb++] = U;
  }
  int n = 0;
  for (int i = 0; i + 1 < nb; ++i) {
    double const a = br[i], b = br[i + 1];
    int const pieces = (int)ceil((b - a) / 2.0);
    int const np = pieces < 1 ? 1 : pieces;
    for (int q = 0; q < np && n < kNfwMisMaxPanels; ++q) {
      lo[n] = a + (b - a) * q / np;
      hi[n] = q == np - 1 ? b : a + (b - a) * (q + 1) / np;
      ++n;
    }
  }
  return n;
}

// NFW_DSIGMA_MIS::operator()(r, rmis, lnM) (nfw_dsigma_mis.cuh:109-123) over
// a table tab = {ln x [nx], ln xmis [nm], T [nm][nx]} (x fastest):
// 2 r_s delta_c rho exp(T.clamp(ln(r / r_s), ln(rmis / r_s))) 1e-12 with
// delta_c = (200 c^3 / 3) / (ln(1 + c) - c / (1 + c)) and
// r_s = cbrt(3 M / (800 pi rho)) / c.
__device__ inline double nfw_dsigma_mis_tab(const double* tab, int nx, int nm, double r,
                                            double rmis, double lnM, double c, double rho)
{
  double const delta_c = (200.0 * c * c * c / 3.0) / (log(1.0 + c) - c / (1.0 + c));
  double const r_200 = cbrt(3.0 * exp(lnM) / (800.0 * 3.14159265358979323846 * rho));
  double const r_s = r_200 / c;
  double const lu = interp2d_clamp(tab, nx, tab + nx, nm, tab + nx + nm, log(r / r_s),
                                   log(rmis / r_s));
  return 2.0 * r_s * delta_c * rho * exp(lu) * 1e-12;
}

// SIG_MAX::operator() (sig_max.cuh:60-68): the halo Delta-Sigma as the larger
// of the one-halo term sigma_1(r, lnM) and bias(zt, lnM) sigma_2(r, zt), each
// a clamped bilinear lookup in the sample's tables (deltasigma/{r_sigma_
// deltasigma, lnM, sigma_1, sigma_2, bias}, matter_power_lin/z; make_interp_
// 2d.cuh:26-40).  Packed as
//   par = {n_r, n_m, n_z, r[n_r], lnM[n_m], z[n_z],
//          sigma_1[n_m][n_r], sigma_2[n_z][n_r], bias[n_m][n_z]}
// (each table x-fastest in its (x, y) = (r, lnM), (r, z), (z, lnM)).
__device__ inline double sig_max(const double* par, double r, double lnM, double zt)
{
  int const nr = int(par[0]), nm = int(par[1]), nz = int(par[2]);
  const double* xr = par + 3;
  const double* xm = xr + nr;
  const double* xz = xm + nm;
  const double* s1 = xz + nz;
  const double* s2 = s1 + (size_t)nm * nr;
  const double* bi = s2 + (size_t)nz * nr;
  // one clamped search per axis, shared by the tables over it
  double const rc = r < xr[0] ? xr[0] : (r > xr[nr - 1] ? xr[nr - 1] : r);
  double const mc = lnM < xm[0] ? xm[0] : (lnM > xm[nm - 1] ? xm[nm - 1] : lnM);
  double const zc = zt < xz[0] ? xz[0] : (zt > xz[nz - 1] ? xz[nz - 1] : zt);
  int const ir = bracket(xr, nr, rc), im = bracket(xm, nm, mc), iz = bracket(xz, nz, zc);
  double const tr = (rc - xr[ir]) / (xr[ir + 1] - xr[ir]);
  double const tm = (mc - xm[im]) / (xm[im + 1] - xm[im]);
  double const tz = (zc - xz[iz]) / (xz[iz + 1] - xz[iz]);
  auto bil = [](const double* z, int nx, int i, int j, double t, double u) {
    const double* r0 = z + j * nx + i;
    const double* r1 = r0 + nx;
    return (1.0 - t) * (1.0 - u) * r0[0] + t * (1.0 - u) * r0[1] + (1.0 - t) * u * r1[0] +
           t * u * r1[1];
  };
  double const a = bil(s1, nr, ir, im, tr, tm);
  double const b = bil(bi, nz, iz, im, tz, tm) * bil(s2, nr, ir, iz, tr, tz);
  return a < b ? b : a;  // std::max(sig_1, sig_2)
}

// The point's DES data (y3g_des_upload): {roffset_tau, LO_LC alpha, a, b,
// R_lambda, MOR_DES A, B, alpha, sigma_intr, epsilon, z_pivot, SIG_MAX
// tables as packed for sig_max()}.
constexpr int kDesLoLc = 1, kDesMor = 5, kDesSig = 11;

// The reference's real-data integrands, FP64 (each the operator() of its
// module), in the volume's own variables except that the two miscentred ones
// take the lambda_obs integral in closed form (lo_lc_bin over the volume's
// lo range, kept in ig.lo_lo / ig.lo_hi): lo enters them only through
// lo_lc, a Gaussian in lo / lc whose width vanishes as rmis -> 0, so left
// to the cubature it is a ridge along lo = mu lc that no axis-aligned split
// resolves.
//   DSIGMA6D_DES  (lc, zt, lnM, rmis, theta)
//     roffset(rmis) lo_lc_bin(lc, rmis) mor_des(lc, lnM, zt) dV/dz hmf Omega
//     / 2 / 3.1415926535897 (the reference's constant) x sig_max(R', lnM, zt)
//     int_zo_zt(zo_lo, zo_hi, zt), R' = sqrt(R^2 + rmis^2 + 2 rmis R cos theta)
//     (sigma_miscent_y1_mort_cuda_scalarintegrand.cu:161-182, its lo axis
//     integrated);
//   DSIGMA3D_DES  (lo, zt, lnM): mor_des(lo, lnM, zt) dV/dz hmf Omega
//     sig_max(R, lnM, zt) int_zo_zt (sigma_cent_y1_mort_cuda_scalarintegrand.cu
//     :155-166);
//   COUNT3D_DES   (lo, zt, lnM): mor_des(lo, lnM, zt) dV/dz hmf Omega int_zo_zt
//     (NC_cent_y1_mort_cuda_scalarintegrand.cu:165-175);
//   COUNT5D_DES   (lc, zt, lnM, rmis): roffset lo_lc_bin mor_des(lc, ...) dV/dz
//     hmf Omega int_zo_zt (NC_miscent_y1_mort_cuda_scalarintegrand.cu:153-164,
//     its lo axis integrated).
//   COUNT4D_CARD  (ln lo, ln lt, zt, lnM): int_lc_lt(lo, lt, zt) int_zo_zt
//     Omega dV/dz hmf mor_des_log(lt, lnM, zt), lo = e^(ln lo), lt = e^(ln lt)
//     (gt_card_gpu/numberCounts.cu:160-174).
// ------------------------------------------ Costanzi-2026 P[X] operator
// P_operator_gpu<W>::operator()(lt, zt, lnM) (p_operator_gpu_t.cuh:172-255)
// and its pieces, FP64.  The point's data (y3g_des_upload) from kDesSig on:
//   {n_chi, n_bm, n_bz, n_xr, n_xz, chi_z, chi_dc, bias_lnM, bias_z,
//    bias[n_bz][n_bm], xi_r, xi_z, xi[n_xz][n_xr]},
// h0 in des[0] and MOR_SHIFTED_POISSON_t's {log10_Mmin, log10_M1, alpha,
// sigma_intr, epsilon, z_pivot} at kDesMor.
#define Y3G_PHOTOZ_TABLE(name, n) static __constant__ double name[n]
#include "y3g_photoz_table.inc"
#undef Y3G_PHOTOZ_TABLE

// SIGMA_PHOTOZ_TABLE_t (sigma_photoz_table_t.cuh:15-75): linear in a
// 120-point uniform table over z in [0.1, 0.9], clamped.
__device__ inline double sigma_photoz_table(double z)
{
  if (z <= 0.1) return kPhotozSigma[0];
  if (z >= 0.9) return kPhotozSigma[119];
  double const dz = (0.9 - 0.1) / 119;
  double const u = (z - 0.1) / dz;
  int const i = int(u);
  int const lo = i < 0 ? 0 : (i >= 119 ? 118 : i);
  double const f = u - lo;
  return kPhotozSigma[lo] + f * (kPhotozSigma[lo + 1] - kPhotozSigma[lo]);
}

// MOR_SHIFTED_POISSON_t::operator()(ltr, lnM, z)
// (mor_shifted_poisson_t.cuh:107-138): l_sat = ((M - Mmin) / (M1 - Mmin))^alpha
// ((1 + z) / (1 + z_p))^eps, mi = (l_sat sigma)^2, lam = l_sat + mi,
// pdf = exp(-lam + (ltr + mi - 1) ln lam - lgamma(ltr + mi)); 0 below Mmin.
__host__ __device__ inline double mor_shifted_poisson(const double* q, double ltr, double lnM,
                                                      double z)
{
  double const Mmin = pow(10.0, q[0]), M1 = pow(10.0, q[1]);
  double const dM = exp(lnM) - Mmin, dM1 = M1 - Mmin;
  if (dM <= 0.0 || dM1 <= 0.0) return 0.0;
  double const l_sat = pow(dM / dM1, q[2]) * pow((1.0 + z) / (1.0 + q[5]), q[4]);
  double const mi = (l_sat * q[3]) * (l_sat * q[3]);
  double const lam = fmax(l_sat + mi, 1.0e-300);
  double const x = ltr + mi;
  if (ltr < 0.0 || x <= 0.0) return 0.0;
  return exp(-lam + (x - 1.0) * log(lam) - lgamma(x));
}

// Two-disk overlap fraction f_A(theta; theta_lob, theta_lt)
// (p_operator_gpu_t.cuh:63-92): the area of the disk of radius theta_lt,
// centred theta away, that lies inside the disk of radius theta_lob, over
// the first disk's area.
__host__ __device__ inline double pop_area_overlap(double th, double lob, double lt)
{
  if (th > lob + lt) return 0.0;
  double const diff = fabs(lob - lt);
  if (th <= diff) {
    if (lt <= lob) return 1.0;
    double const r = lob / lt;
    return r * r;
  }
  double c1 = (th * th + lt * lt - lob * lob) / (2.0 * th * lt);
  double c2 = (th * th + lob * lob - lt * lt) / (2.0 * th * lob);
  c1 = fmax(-1.0, fmin(1.0, c1));
  c2 = fmax(-1.0, fmin(1.0, c2));
  double const sq = fmax((-th + lt + lob) * (th + lt - lob) * (th - lt + lob) * (th + lt + lob),
                         0.0);
  double const A = lt * lt * acos(c1) + lob * lob * acos(c2) - 0.5 * sqrt(sq);
  return A / (3.14159265358979323846 * lt * lt);
}

// Exclusion angle at z (p_operator_gpu_t.cuh:95-104): the law of cosines
// for R_excl between comoving distances chi_z and chi_o.
__host__ __device__ inline double pop_theta_excl(double chi_z, double chi_o, double R_excl)
{
  double const den = 2.0 * chi_z * chi_o;
  if (den < 1e-30) return 0.0;
  double c = (chi_z * chi_z + chi_o * chi_o - R_excl * R_excl) / den;
  c = fmax(-1.0, fmin(1.0, c));
  if (c >= 1.0 - 1e-12) return 1e-6;
  return acos(c);
}

template <int K>
__device__ inline double integrand_pop_k(const y3g_integral& ig, const PointDev& P,
                                         Tabs<double> tb, const double* des, const double* y)
{
  double const lt = y[0], zt = y[1], lnM = y[2];
  double const lobc = ig.lo_hi;  // the bin's centre, the upper lt limit
  if (lt <= 0.0 || lt > lobc) return 0.0;
  double const zob = 0.5 * (ig.zo_lo + ig.zo_hi);
  double const sz = sigma_photoz_table(zt);
  double const u = (zt - zob) / sz;
  if (fabs(u) >= 1.0) return 0.0;
  double const wz = 1.0 - u * u;
  double const common = hmf<double>(P, tb.hmf, lnM, zt) * dv_do_dz<double>(P, tb.da, zt) *
                        mor_shifted_poisson(des + kDesMor, lt, lnM, zt);
  if (common <= 0.0) return 0.0;
  const double* q = des + kDesSig;
  int const nc = int(q[0]), nbm = int(q[1]), nbz = int(q[2]), nxr = int(q[3]), nxz = int(q[4]);
  const double* chz = q + 5;
  const double* chd = chz + nc;
  const double* bm = chd + nc;
  const double* bz = bm + nbm;
  const double* bt = bz + nbz;
  const double* xr = bt + nbm * nbz;
  const double* xz = xr + nxr;
  const double* xt = xz + nxz;
  double const h0 = des[0];
  double const lob = lobc;
  double const chi_o = emg_des_curve(chz, chd, nc, zob) * h0;
  double const chi_z = emg_des_curve(chz, chd, nc, zt) * h0;
  double const rl_lob = pow(lob / 100.0, 0.2);
  double const theta_lob = rl_lob * (1.0 + zob) / chi_o;
  double const R_excl = rl_lob * (1.0 + zob);
  double const th_excl = pop_theta_excl(chi_z, chi_o, R_excl);
  double const th_hi = 2.0 * theta_lob;
  if (th_excl >= th_hi) return 0.0;
  double const theta_lt = pow(lt / 100.0, 0.2) * (1.0 + zt) / chi_z;
  double const k_sig = 2.5 / theta_lob;
  double const th0 = 0.5 * theta_lob;
  constexpr double gx[5] = {-0.9061798459386640, -0.5384693101056831, 0.0,
                            0.5384693101056831, 0.9061798459386640};
  constexpr double gw[5] = {0.2369268850561891, 0.4786286704993665, 0.5688888888888889,
                            0.4786286704993665, 0.2369268850561891};
  double const th_mid = fmax(th_excl, fmin(theta_lob, th_hi));
  double th_sum = 0.0;
  for (int panel = 0; panel < 2; ++panel) {
    double const a = panel == 0 ? th_excl : th_mid, b = panel == 0 ? th_mid : th_hi;
    if (!(b > a + 1e-10)) continue;
    double const mid = 0.5 * (a + b), hl = 0.5 * (b - a);
    for (int it = 0; it < 5; ++it) {
      double const th = mid + hl * gx[it];
      double const fA = pop_area_overlap(th, theta_lob, theta_lt);
      if (fA <= 0.0) continue;
      double const wth = hl * gw[it] * 2.0 * 3.14159265358979323846 * sin(th);
      double F = 1.0;
      if constexpr (K != Y3G_POP_P1) {
        double const d3 =
          sqrt(fmax(chi_z * chi_z + chi_o * chi_o - 2.0 * chi_z * chi_o * cos(th), 0.0));
        double const xi = interp2d_clamp(xr, nxr, xz, nxz, xt, d3, zob);
        double const sg = 1.0 / (1.0 + exp(-k_sig * (th - th0)));
        F = K == Y3G_POP_I1 ? xi * sg : K == Y3G_POP_I2 ? xi : xi * (1.0 - sg);
      }
      th_sum += wth * fA * F;
    }
  }
  double const b_mz = K == Y3G_POP_P1 ? 1.0 : interp2d_clamp(bm, nbm, bz, nbz, bt, lnM, zt);
  return wz * lt * common * b_mz * th_sum;
}

// Not inlined: inlined into the cubature's node loop it raised the loop's
// spills (237 VGPRs spilled against 161) and ran at half the speed.
template <int K>
__device__ __noinline__ double integrand_des_k(const y3g_integral& ig, const PointDev& P,
                                         Tabs<double> tb, const double* des, const double* y,
                                         MorDesTabs const& mt = mor_des_const_tabs())
{
  if constexpr (is_pop_kind(K)) return integrand_pop_k<K>(ig, P, tb, des, y);
  constexpr int iz = KindAxes<K>::iz, im = KindAxes<K>::im;
  if constexpr (K == Y3G_COUNT4D_CARD) {
    double const zt = y[iz], lnM = y[im];
    double const lt = exp(y[1]), lo = exp(y[0]);
    double const mor_v = mor_des_log(des + kDesMor, lt, lnM, zt, mt);
    double const common = omega_z_des(zt) * dv_do_dz<double>(P, tb.da, zt) *
                          hmf<double>(P, tb.hmf, lnM, zt) * mor_v;
    return int_lc_lt_des(des + kDesSig, lo, lt, zt) *
           int_zo_zt_des<double>(ig.zo_lo, ig.zo_hi, zt) * common;
  }
  double const zt = y[iz], lnM = y[im];
  constexpr bool miscent = K == Y3G_DSIGMA6D_DES || K == Y3G_COUNT5D_DES;
  double const lam = y[0];  // the MOR_DES richness argument: lc, or lo (centred)
  double common = mor_des(des + kDesMor, lam, lnM, zt, mt) * dv_do_dz<double>(P, tb.da, zt) *
                  hmf<double>(P, tb.hmf, lnM, zt) * omega_z_des(zt);
  if constexpr (miscent) {
    double const rmis = y[3];
    common = roffset_gamma<double>(rmis, des[0]) *
             lo_lc_bin<double>(ig.lo_lo, ig.lo_hi, y[0], rmis, des[kDesLoLc],
                               des[kDesLoLc + 1], des[kDesLoLc + 2], des[kDesLoLc + 3]) *
             common;
  }
  double const izo = int_zo_zt_des<double>(ig.zo_lo, ig.zo_hi, zt);
  if constexpr (K == Y3G_DSIGMA6D_DES) {
    double const R = ig.radius, rmis = y[3];
    double const rs = sqrt(R * R + rmis * rmis + 2 * rmis * R * cos(y[4]));
    return sig_max(des + kDesSig, rs, lnM, zt) * izo * (common / 2.0 / 3.1415926535897);
  } else if constexpr (K == Y3G_DSIGMA3D_DES) {
    return sig_max(des + kDesSig, ig.radius, lnM, zt) * izo * common;
  } else {
    return izo * common;
  }
}

__device__ inline double integrand_des(const y3g_integral& ig, const PointDev& P,
                                       Tabs<double> tb, const double* des, const double* y)
{
  switch (ig.kind) {
    case Y3G_DSIGMA6D_DES: return integrand_des_k<Y3G_DSIGMA6D_DES>(ig, P, tb, des, y);
    case Y3G_DSIGMA3D_DES: return integrand_des_k<Y3G_DSIGMA3D_DES>(ig, P, tb, des, y);
    case Y3G_COUNT3D_DES: return integrand_des_k<Y3G_COUNT3D_DES>(ig, P, tb, des, y);
    case Y3G_COUNT4D_CARD: return integrand_des_k<Y3G_COUNT4D_CARD>(ig, P, tb, des, y);
    case Y3G_POP_P1: return integrand_des_k<Y3G_POP_P1>(ig, P, tb, des, y);
    case Y3G_POP_I1: return integrand_des_k<Y3G_POP_I1>(ig, P, tb, des, y);
    case Y3G_POP_I2: return integrand_des_k<Y3G_POP_I2>(ig, P, tb, des, y);
    case Y3G_POP_J: return integrand_des_k<Y3G_POP_J>(ig, P, tb, des, y);
    default: return integrand_des_k<Y3G_COUNT5D_DES>(ig, P, tb, des, y);
  }
}

// ------------------------------------------------ Genz-Malik rule nodes
constexpr double kL2 = 0.35856858280031809;  // sqrt(9/70)
constexpr double kL3 = 0.94868329805051380;  // sqrt(9/10)
constexpr double kL5 = 0.68824720161168529;  // sqrt(9/19)

__host__ __device__ constexpr int gm_num_nodes(int d)
{
  return 1 + 4 * d + 2 * d * (d - 1) + (1 << d);
}

// Generator class (1..5) of node j, and its unit offset on axis a.  Node
// enumeration: [0] centre; +-l2 e_i; +-l3 e_i; +-l4 e_i +-l4 e_k (i<k);
// the 2^d +-l5 corners.
Y3G_HD int gm_class(int d, int j)
{
  if (j == 0) return 1;
  j -= 1;
  if (j < 2 * d) return 2;
  j -= 2 * d;
  if (j < 2 * d) return 3;
  j -= 2 * d;
  if (j < 2 * d * (d - 1)) return 4;
  return 5;
}

// Unit offset of node j on axis a (no local arrays: the node coordinates
// stay in registers).
Y3G_HD double gm_offset(int d, int j, int a)
{
  if (j == 0) return 0.0;
  j -= 1;
  if (j < 2 * d) return (j >> 1) == a ? ((j & 1) ? -kL2 : kL2) : 0.0;
  j -= 2 * d;
  if (j < 2 * d) return (j >> 1) == a ? ((j & 1) ? -kL3 : kL3) : 0.0;
  j -= 2 * d;
  if (j < 2 * d * (d - 1)) {
    int const pair = j >> 2, sg = j & 3;
    // pair -> (p, q), p < q, lexicographic
    int p = 0, cnt = 0;
    while (pair >= cnt + (d - 1 - p)) cnt += d - 1 - p, ++p;
    int const q = p + 1 + (pair - cnt);
    if (a == p) return (sg & 1) ? -kL3 : kL3;
    if (a == q) return (sg & 2) ? -kL3 : kL3;
    return 0.0;
  }
  j -= 2 * d * (d - 1);
  return ((j >> a) & 1) ? -kL5 : kL5;
}

// Weights per generator class (1..5) of the degree-7 rule, its embedded
// degree-5 rule, and a degree-3 rule on the centre and the +-l3 axis nodes
// (w3 = 1 / (6 l3^2) on those, 1 - 2d w3 at the centre).
struct GmWeights {
  double w7[6], w5[6], w3[6];
};
__host__ __device__ constexpr GmWeights gm_weights_of(int d)
{
  GmWeights w{};
  double const dd = d;
  w.w3[3] = 10.0 / 54.0;
  w.w3[1] = 1.0 - 2.0 * dd * w.w3[3];
  w.w7[1] = (12824.0 - 9120.0 * dd + 400.0 * dd * dd) / 19683.0;
  w.w7[2] = 980.0 / 6561.0;
  w.w7[3] = (1820.0 - 400.0 * dd) / 19683.0;
  w.w7[4] = 200.0 / 19683.0;
  w.w7[5] = 6859.0 / 19683.0 / double(1 << d);
  w.w5[1] = (729.0 - 950.0 * dd + 50.0 * dd * dd) / 729.0;
  w.w5[2] = 245.0 / 486.0;
  w.w5[3] = (265.0 - 100.0 * dd) / 1458.0;
  w.w5[4] = 25.0 / 729.0;
  return w;
}
__host__ __device__ inline void gm_weights(int d, double* w7, double* w5, double* w3)
{
  GmWeights const w = gm_weights_of(d);
  for (int g = 0; g < 6; ++g) w7[g] = w.w7[g], w5[g] = w.w5[g], w3[g] = w.w3[g];
}

// --------------------------------------------------------- Philox4x32-10
Y3G_HD void philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t const hi0 = __umulhi(0xD2511F53u, c[0]);
    uint32_t const lo0 = 0xD2511F53u * c[0];
    uint32_t const hi1 = __umulhi(0xCD9E8D57u, c[2]);
    uint32_t const lo1 = 0xCD9E8D57u * c[2];
    uint32_t const n0 = hi1 ^ c[1] ^ k0;
    uint32_t const n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

}  // namespace y3g
