/*
 * y3o_pop.c -- FP64 restatement of the Costanzi-2026 P[X] operator GPU
 * integrands P_operator_gpu<W> (src/models/p_operator_gpu_t.cuh:118-299) and
 * their pieces.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): linked by tests/ and never by the
 * product library.
 *
 * Pins (tests/test_oracle_pop.py):
 *   y3o_mor_shifted_poisson  mor_shifted_poisson_t.cuh:107-138   the Poisson pmf at
 *                            integer ltr + mi - 1 (scipy.stats.poisson); 0 below
 *                            Mmin
 *   y3o_sigma_photoz_table   sigma_photoz_table_t.cuh:15-75      table values at the
 *                            knots read by hand; linear between; clamps
 *   y3o_pop_area_overlap     p_operator_gpu_t.cuh:63-92          the lens area of two
 *                            disks by quadrature of chord lengths; the
 *                            containment and disjoint cases
 *   y3o_pop_theta_excl       p_operator_gpu_t.cuh:95-104         law of cosines
 *   y3o_integrand_pop        p_operator_gpu_t.cuh:172-255        J + I1 = I2; I2 = b xi
 *                            P1 for constant bias and xi_NL tables; zero outside
 *                            lt in (0, lob], |zt - zob| < sigma_z; the angle sum
 *                            against adaptive quadrature of the same angle
 *                            integrand
 */
#include "y3o.h"

#include <math.h>

#define POP_PI 3.14159265358979323846

#define Y3G_PHOTOZ_TABLE(name, n) static const double name[n]
#include "../y3_cluster_cpp_amd/csrc/y3g_photoz_table.inc"
#undef Y3G_PHOTOZ_TABLE

double
y3o_sigma_photoz_table(double z)
{
  if (z <= 0.1) return kPhotozSigma[0];
  if (z >= 0.9) return kPhotozSigma[119];
  double const dz = (0.9 - 0.1) / 119;
  double const u = (z - 0.1) / dz;
  int const i = (int)u;
  int const lo = i < 0 ? 0 : (i >= 119 ? 118 : i);
  double const f = u - lo;
  return kPhotozSigma[lo] + f * (kPhotozSigma[lo + 1] - kPhotozSigma[lo]);
}

/* q = {log10_Mmin, log10_M1, alpha, sigma_intr, epsilon, z_pivot} */
double
y3o_mor_shifted_poisson(const double* q, double ltr, double lnM, double z)
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

double
y3o_pop_area_overlap(double th, double lob, double lt)
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
  double const sq =
    fmax((-th + lt + lob) * (th + lt - lob) * (th - lt + lob) * (th + lt + lob), 0.0);
  double const A = lt * lt * acos(c1) + lob * lob * acos(c2) - 0.5 * sqrt(sq);
  return A / (POP_PI * lt * lt);
}

double
y3o_pop_theta_excl(double chi_z, double chi_o, double R_excl)
{
  double const den = 2.0 * chi_z * chi_o;
  if (den < 1e-30) return 0.0;
  double c = (chi_z * chi_z + chi_o * chi_o - R_excl * R_excl) / den;
  c = fmax(-1.0, fmin(1.0, c));
  if (c >= 1.0 - 1e-12) return 1e-6;
  return acos(c);
}

/* The angle integrand of P_operator_gpu: 2 pi sin(theta) f_A F(theta), with
 * F = 1 (P1), b xi sigma (I1), b xi (I2), b xi (1 - sigma) (J) before the
 * bias factor; exposed for the angle-sum pin. */
double
y3o_pop_theta_term(int kind, double th, double theta_lob, double theta_lt, double chi_z,
                   double chi_o, double zob, const double* xr, int nxr, const double* xz,
                   int nxz, const double* xt)
{
  double const fA = y3o_pop_area_overlap(th, theta_lob, theta_lt);
  if (fA <= 0.0) return 0.0;
  double F = 1.0;
  if (kind != 8) {
    double const d3 = sqrt(fmax(chi_z * chi_z + chi_o * chi_o - 2.0 * chi_z * chi_o * cos(th),
                                0.0));
    double const xi = y3o_interp2d_clamp(xr, nxr, xz, nxz, xt, d3, zob);
    double const sg = 1.0 / (1.0 + exp(-(2.5 / theta_lob) * (th - 0.5 * theta_lob)));
    F = kind == 9 ? xi * sg : kind == 10 ? xi : xi * (1.0 - sg);
  }
  return 2.0 * POP_PI * sin(th) * fA * F;
}

/* P_operator_gpu<W>::operator()(lt, zt, lnM) (p_operator_gpu_t.cuh:172-255),
 * kinds 8 P1, 9 I1, 10 I2, 11 J; des as y3g_des_upload's P[X] layout; the
 * bin centre in ig->lo_hi. */
double
y3o_integrand_pop(const y3o_params* p, const y3o_tables* t, const double* des,
                  const y3o_integral* ig, const double* x)
{
  double const lt = x[0], zt = x[1], lnM = x[2];
  double const lobc = ig->lo_hi;
  if (lt <= 0.0 || lt > lobc) return 0.0;
  double const zob = 0.5 * (ig->zo_lo + ig->zo_hi);
  double const u = (zt - zob) / y3o_sigma_photoz_table(zt);
  if (fabs(u) >= 1.0) return 0.0;
  double const wz = 1.0 - u * u;
  double const common = y3o_hmf(p, t, lnM, zt) * y3o_dvdz(p, t, zt) *
                        y3o_mor_shifted_poisson(des + 5, lt, lnM, zt);
  if (common <= 0.0) return 0.0;
  const double* q = des + 11;
  int const nc = (int)q[0], nbm = (int)q[1], nbz = (int)q[2], nxr = (int)q[3], nxz = (int)q[4];
  const double* chz = q + 5;
  const double* chd = chz + nc;
  const double* bm = chd + nc;
  const double* bz = bm + nbm;
  const double* bt = bz + nbz;
  const double* xr = bt + nbm * nbz;
  const double* xz = xr + nxr;
  const double* xt = xz + nxz;
  double const h0 = des[0];
  double const chi_o = y3o_interp1d_clamp(chz, chd, nc, zob) * h0;
  double const chi_z = y3o_interp1d_clamp(chz, chd, nc, zt) * h0;
  double const rl = pow(lobc / 100.0, 0.2);
  double const theta_lob = rl * (1.0 + zob) / chi_o;
  double const th_excl = y3o_pop_theta_excl(chi_z, chi_o, rl * (1.0 + zob));
  double const th_hi = 2.0 * theta_lob;
  if (th_excl >= th_hi) return 0.0;
  double const theta_lt = pow(lt / 100.0, 0.2) * (1.0 + zt) / chi_z;
  static const double gx[5] = {-0.9061798459386640, -0.5384693101056831, 0.0,
                               0.5384693101056831, 0.9061798459386640};
  static const double gw[5] = {0.2369268850561891, 0.4786286704993665, 0.5688888888888889,
                               0.4786286704993665, 0.2369268850561891};
  double const th_mid = fmax(th_excl, fmin(theta_lob, th_hi));
  double th_sum = 0.0;
  for (int panel = 0; panel < 2; ++panel) {
    double const a = panel == 0 ? th_excl : th_mid, b = panel == 0 ? th_mid : th_hi;
    if (!(b > a + 1e-10)) continue;
    double const mid = 0.5 * (a + b), hl = 0.5 * (b - a);
    for (int it = 0; it < 5; ++it)
      th_sum += hl * gw[it] *
                y3o_pop_theta_term(ig->kind, mid + hl * gx[it], theta_lob, theta_lt, chi_z,
                                   chi_o, zob, xr, nxr, xz, nxz, xt);
  }
  double const b_mz = ig->kind == 8 ? 1.0 : y3o_interp2d_clamp(bm, nbm, bz, nbz, bt, lnM, zt);
  return wz * lt * common * b_mz * th_sum;
}
