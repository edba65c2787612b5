/*
 * y3o_selgl.c -- FP64 CPU restatement of the reference's fixed
 * Gauss-Legendre number-count evaluator (NumCountsSelGL over SelGLCore,
 * src/models/n_operator_sel_gl_t.hh:72-125, 171-201), with the selection
 * function of src/models/sel_function_t.hh:17-60 and the effective survey
 * area of src/models/omega_z_des.hh:15-33.
 *
 * TEST INFRASTRUCTURE ONLY (see y3o.h).
 *
 * Pins (tests/test_oracle_selgl.py):
 *   y3o_omega_z_des      values at points of each polynomial piece computed
 *                        by hand from the reference coefficients (a reversed
 *                        coefficient order or a wrong piece boundary fails);
 *   y3o_sel_function     zero outside the axes, bilinear inside: exact on a
 *                        bilinear table, knot values;
 *   y3o_gl_nodes_ab      exactness on degree 2n-1 monomials over [a, b];
 *   y3o_selgl_counts     constant integrand (closed form: area x constant),
 *                        separable integrand (product of two 1-D GL sums),
 *                        and convergence, as the node counts grow, to an
 *                        adaptive nested Gauss-Kronrod integral of the same
 *                        integrand.
 */
#include <math.h>
#include <stdlib.h>

#include "y3o.h"

/* omega_z_des.hh:15-33: three degree-5 polynomials, coefficients listed
 * leading power first (utils/polynomial.hh Horner order); the middle piece
 * is evaluated at z - 0.6. */
double
y3o_omega_z_des(double z)
{
  static const double c1[6] = {0.0, 0.0, 0.0, -0.00262353, 0.01940118, 0.45133063};
  static const double c2[6] = {1.33647377e+4, 1.35291046e+3, -1.26204891e+2,
                               -2.83454918e+1, -2.26465905, 3.84958753e-1};
  static const double c3[6] = {0, 0, -1.88101967, 4.8071839, -4.11424324, 1.18196785};
  const double* c = z < 0.504 ? c1 : (z < 0.7 ? c2 : c3);
  double const x = (z >= 0.504 && z < 0.7) ? z - 0.6 : z;
  double out = 0.0;
  for (int i = 0; i < 6; ++i) out = c[i] + x * out;
  return out;
}

/* sel_function_t.hh:28-35: zero outside [lnm_min, lnm_max] x [z_min, z_max],
 * else the clamped bilinear interpolant of the bin's (n_z, n_lnm) plane
 * (lnM fastest, interp_2d.hh). */
double
y3o_sel_function(const double* lnm, int nm, const double* z, int nz, const double* plane,
                 double lnM, double zt)
{
  if (lnM < lnm[0] || lnM > lnm[nm - 1] || zt < z[0] || zt > z[nz - 1]) return 0.0;
  return y3o_interp2d_clamp(lnm, nm, z, nz, plane, lnM, zt);
}

/* p_operator_cuhre_t gl_nodes: n Gauss-Legendre nodes and weights on [a, b]. */
void
y3o_gl_nodes_ab(int n, double a, double b, double* x, double* w)
{
  y3o_gl_rule(n, x, w);
  double const c = 0.5 * (a + b), h = 0.5 * (b - a);
  for (int i = 0; i < n; ++i) {
    x[i] = c + h * x[i];
    w[i] *= h;
  }
}

/* n_operator_sel_gl_t.hh:88-118 (SelGLCore::build_weights, number counts):
 *   zfac_q = w_q dV/dz(z_q) Omega(z_q)
 *   W(lnM_k) = sum_q zfac_q hmf(lnM_k, z_q) S(lnM_k, z_q)
 *   N = sum_k w_k W(lnM_k)     (NumCountsSelGL::evaluate returns this norm)
 * hmf and dV/dz from the point's tables (y3o_hmf, y3o_dvdz). */
double
y3o_selgl_counts(const y3o_params* p, const y3o_tables* t, const double* lnm, int nm,
                 const double* z, int nz, const double* plane, int n_lnm, int n_z,
                 double lnm_lo, double lnm_hi, double zt_lo, double zt_hi)
{
  double* xm = (double*)malloc(sizeof(double) * 2 * (size_t)(n_lnm + n_z));
  if (!xm) return NAN;
  double* wm = xm + n_lnm;
  double* xz = wm + n_lnm;
  double* wz = xz + n_z;
  y3o_gl_nodes_ab(n_lnm, lnm_lo, lnm_hi, xm, wm);
  y3o_gl_nodes_ab(n_z, zt_lo, zt_hi, xz, wz);
  double n0 = 0.0;
  for (int k = 0; k < n_lnm; ++k) {
    double acc = 0.0;
    for (int q = 0; q < n_z; ++q) {
      double const zfac = wz[q] * y3o_dvdz(p, t, xz[q]) * y3o_omega_z_des(xz[q]);
      acc += zfac * y3o_hmf(p, t, xm[k], xz[q]) *
             y3o_sel_function(lnm, nm, z, nz, plane, xm[k], xz[q]);
    }
    n0 += wm[k] * acc;
  }
  free(xm);
  return n0;
}
