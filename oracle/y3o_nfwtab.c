/*
 * y3o_nfwtab.c -- FP64 restatement of the tabulated miscentred NFW profile
 * NFW_DSIGMA_MIS (src/models/nfw_dsigma_mis.cuh:51-131) and of the table it
 * reads (data/nfw_off_center/table_1000_1e-03_5e+03_log_deltasigma_gamma.txt,
 * absent from the reference tree; its axes are the ..._gamma_logx.txt and
 * ..._gamma_logxmis.txt files next to it).
 *
 * TEST INFRASTRUCTURE ONLY (the checker): linked by tests/ and never by the
 * product library.
 *
 * The table holds T(ln x, ln xmis) = ln(G(x, xmis) / 2), where
 *   G(x, m) = int_0^inf u e^-u Ring(x, m u) du     (gamma offset kernel,
 *             roffset_t.hh:18-23, with scale m = xmis)
 *   Ring(x, s) = (1/pi) int_0^pi g(X) cos 2 beta dphi
 *             (tangential Delta-Sigma of an NFW halo offset by s, averaged on
 *             the ring of radius x; X^2 = (x - s)^2 + 4 x s sin^2(phi/2),
 *             cos beta = (x - s cos phi) / X, phi = pi - theta)
 * in units of r_s (x = r / r_s, xmis = rmis / r_s) and of rho_s r_s
 * (g = y3o_nfw_g), so that NFW_DSIGMA_MIS = 2 r_s delta_c rho e^T 1e-12.
 *
 * Pins (tests/test_oracle_nfwtab.py):
 *   y3o_nfw_mis_unit      the adaptive y3o_dsigma_mis_gamma (independent
 *                         nested GK21 quadrature) at 1e-7 over the table's
 *                         (x, xmis) range
 *   y3o_nfw_mis_axes      data/nfw_off_center/table_..._gamma_logx.txt and
 *                         table_..._gamma_logxmis.txt
 *   y3o_nfw_dsigma_mis_lookup  the 30 cluster_toolkit values of
 *                         test/nfw_dsigma_mis.test.cu:32-66 at 5e-3 (through
 *                         the table's knots around each point)
 */
#include "y3o.h"

#include <math.h>

#define NT_PI 3.14159265358979323846

/* 8-point Gauss-Legendre on [-1, 1], positive half. */
static const double g8x[4] = {0.1834346424956498, 0.5255324099163290,
                              0.7966664774136267, 0.9602898564975363};
static const double g8w[4] = {0.3626837833783620, 0.3137066458778873,
                              0.2223810344533745, 0.1012285362903763};

/* Ring(x, s): phi in [0, pi] on panels graded geometrically toward phi = 0
 * (theta = pi, where the offset centre comes closest): breakpoints
 * pi 2^(j - L), j = 0 .. L, with L chosen so that the first panel is at most
 * 1/16 of the feature width phi_c = |x - s| / sqrt(x s); 8-point GL each. */
int
y3o_nfw_ring_levels(double x, double s)
{
  double const pc = fabs(x - s) / sqrt(x * s);
  int L = 1;
  if (pc < NT_PI) {
    double const q = log2(NT_PI / fmax(pc, 1e-300));
    L = q > 36.0 ? 40 : (int)ceil(q) + 4;
  }
  if (L > 40) L = 40;
  return L;
}

static double nfw_ring_sum(double x, double s, int absval);

double
y3o_nfw_ring(double x, double s)
{
  return nfw_ring_sum(x, s, 0);
}

/* The ring sum of |g(X) cos 2 beta|: the scale of the terms that Ring(x, s)
 * adds up (their cancellation sets the rounding of a reordered sum). */
static double
nfw_ring_sum(double x, double s, int absval)
{
  int const L = y3o_nfw_ring_levels(x, s);
  double acc = 0.0, a = 0.0;
  for (int j = 0; j <= L; ++j) {
    double const b = NT_PI * ldexp(1.0, j - L);
    double const c = 0.5 * (a + b), h = 0.5 * (b - a);
    double ps = 0.0;
    for (int k = 0; k < 8; ++k) {
      double const t = k < 4 ? -g8x[3 - k] : g8x[k - 4];
      double const w = k < 4 ? g8w[3 - k] : g8w[k - 4];
      double const phi = c + h * t;
      double const sn = sin(0.5 * phi);
      double const d = x - s;
      double const X2 = d * d + 4.0 * x * s * sn * sn;
      double const X = sqrt(X2);
      double const cb = d + 2.0 * s * sn * sn; /* x - s cos(phi) */
      double const c2b = 2.0 * cb * cb / X2 - 1.0;
      double const term = w * y3o_nfw_g(X) * c2b;
      ps += absval ? fabs(term) : term;
    }
    acc += h * ps;
    a = b;
  }
  return acc / NT_PI;
}

/* Outer breakpoints of G(x, m) on u in [0, 60]: graded geometrically toward
 * u* = x / m from both sides (24 levels each), every panel then cut into
 * equal pieces no wider than 2; returns the number of panels written to
 * (lo[i], hi[i]), at most Y3O_NT_MAXPAN. */
int
y3o_nfw_mis_panels(double x, double m, double* lo, double* hi)
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
    for (int q = 0; q < np && n < Y3O_NT_MAXPAN; ++q) {
      lo[n] = a + (b - a) * q / np;
      hi[n] = q == np - 1 ? b : a + (b - a) * (q + 1) / np;
      ++n;
    }
  }
  return n;
}

/* G(x, m) = int_0^60 u e^-u Ring(x, m u) du on those panels, 8-point GL. */
double
y3o_nfw_mis_unit(double x, double m)
{
  double lo[Y3O_NT_MAXPAN], hi[Y3O_NT_MAXPAN];
  int const n = y3o_nfw_mis_panels(x, m, lo, hi);
  double acc = 0.0;
  for (int i = 0; i < n; ++i) {
    double const c = 0.5 * (lo[i] + hi[i]), h = 0.5 * (hi[i] - lo[i]);
    double ps = 0.0;
    for (int k = 0; k < 8; ++k) {
      double const t = k < 4 ? -g8x[3 - k] : g8x[k - 4];
      double const w = k < 4 ? g8w[3 - k] : g8w[k - 4];
      double const u = c + h * t;
      ps += w * u * exp(-u) * y3o_nfw_ring(x, m * u);
    }
    acc += h * ps;
  }
  return acc;
}

/* The same sum over the absolute values of the terms. */
double
y3o_nfw_mis_unit_abs(double x, double m)
{
  double lo[Y3O_NT_MAXPAN], hi[Y3O_NT_MAXPAN];
  int const n = y3o_nfw_mis_panels(x, m, lo, hi);
  double acc = 0.0;
  for (int i = 0; i < n; ++i) {
    double const c = 0.5 * (lo[i] + hi[i]), h = 0.5 * (hi[i] - lo[i]);
    double ps = 0.0;
    for (int k = 0; k < 8; ++k) {
      double const t = k < 4 ? -g8x[3 - k] : g8x[k - 4];
      double const w = k < 4 ? g8w[3 - k] : g8w[k - 4];
      double const u = c + h * t;
      ps += w * u * exp(-u) * nfw_ring_sum(x, m * u, 1);
    }
    acc += h * ps;
  }
  return acc;
}

/* The table's axes: ln x on 1000 points over [ln 1e-3, ln 5e3] and ln xmis on
 * 250 points over [ln 1e-2, ln 5e2], uniform (the reference's file names
 * table_1000_1e-03_5e+03_*; values as in its *_logx.txt / *_logxmis.txt). */
void
y3o_nfw_mis_axes(double* logx, double* logxmis)
{
  double const a = log(1e-3), b = log(5e3), c = log(1e-2), d = log(5e2);
  for (int i = 0; i < Y3O_NT_NX; ++i) logx[i] = a + i * ((b - a) / (Y3O_NT_NX - 1));
  for (int j = 0; j < Y3O_NT_NM; ++j) logxmis[j] = c + j * ((d - c) / (Y3O_NT_NM - 1));
}

/* The table entry T[j][i] = ln(G(x_i, xmis_j) / 2) (x fastest, the order of
 * Interp2D's value vector, interp_2d.cc:26-39). */
double
y3o_nfw_mis_entry(double logx, double logxmis)
{
  return log(0.5 * y3o_nfw_mis_unit(exp(logx), exp(logxmis)));
}

/* NFW_DSIGMA_MIS::operator()(r, rmis, lnM) (nfw_dsigma_mis.cuh:109-123):
 * delta_c = (200 c^3 / 3) / (ln(1 + c) - c / (1 + c)),
 * r_200 = cbrt(3 M / (800 pi rho)), r_s = r_200 / c,
 * 2 r_s delta_c rho exp(T.clamp(ln(r / r_s), ln(rmis / r_s))) 1e-12. */
double
y3o_nfw_dsigma_mis_lookup(const double* logx, int nx, const double* logxmis, int nm,
                          const double* T, double r, double rmis, double lnM,
                          double c, double rho)
{
  double const delta_c = (200.0 * c * c * c / 3.0) / (log(1.0 + c) - c / (1.0 + c));
  double const r_200 = cbrt(3.0 * exp(lnM) / (800.0 * NT_PI * rho));
  double const r_s = r_200 / c;
  double const lu = y3o_interp2d_clamp(logx, nx, logxmis, nm, T, log(r / r_s),
                                       log(rmis / r_s));
  return 2.0 * r_s * delta_c * rho * exp(lu) * 1e-12;
}
