/*
 * asan_check.c -- drives the oracle's code paths under AddressSanitizer and
 * UndefinedBehaviorSanitizer (`make -C oracle asan`, run by
 * tests/test_sanitizers.py).  TEST INFRASTRUCTURE ONLY.
 *
 * Each path runs at a small size: tables of a parameter point, every
 * functor, the mapped integrands, the cubature and VEGAS of all three kinds,
 * the golden integrator, the real-data integrand and its cubature, the fixed
 * Gauss-Legendre counts, the tabulated LC_LT_t and the NFW_DSIGMA_MIS table
 * entries and lookup.  A sanitizer finding aborts the run (the Makefile
 * builds with -fno-sanitize-recover); the values are checked only for being
 * finite.
 */
#include "y3o.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

static int bad = 0;

static void
chk(const char* what, double v)
{
  if (!isfinite(v)) {
    fprintf(stderr, "not finite: %s = %g\n", what, v);
    ++bad;
  }
}

int
main(void)
{
  y3o_params p;
  y3o_default_params(&p);
  y3o_table_spec s;
  y3o_spec_wide(&s);
  y3o_tables t;
  if (y3o_tables_build(&p, &s, &t)) return 2;

  /* functors */
  chk("hmf", y3o_hmf(&p, &t, 33.0, 0.3));
  chk("dvdz", y3o_dvdz(&p, &t, 0.3));
  chk("sci", y3o_sci(&t, 0.3));
  chk("mor", y3o_mor_lognormal(&p, 25.0, 33.0, 0.3));
  chk("lo_lt_pdf", y3o_lo_lt_pdf(&p, 25.0, 24.0));
  chk("lo_lt_bin", y3o_lo_lt_bin(&p, 20.0, 30.0, 24.0));
  chk("zo_bin", y3o_zo_bin(&p, 0.2, 0.35, 0.3));
  chk("lo_lc", y3o_lo_lc(30.0, 32.0, 0.3, 1.66, 0.26, 1.43, 1.0));
  chk("int_zo_zt_des", y3o_int_zo_zt_des(0.2, 0.35, 0.3));
  chk("nfw_dsigma", y3o_nfw_dsigma(1.0, 33.0, 0.3 * 2.77533742639e11, 0.0));
  chk("dsigma_offset", y3o_dsigma_offset(1.0, 0.4, 2.0, 33.0, 0.3 * 2.77533742639e11, 0.0));
  chk("dsigma_mis_gamma",
      y3o_dsigma_mis_gamma(0.5, 0.2, 32.2, 0.3 * 2.77533742639e11, 4.0, 1e-6));
  chk("lc_lt", y3o_lc_lt(30.0, 28.0, 0.25));
  chk("omega_z", y3o_omega_z_des(0.4));

  /* integrals of the three kinds: cubature and VEGAS */
  y3o_cub_opts co;
  memset(&co, 0, sizeof(co));
  co.eps_rel = 1e-3, co.eps_abs = 1e-300, co.max_eval = 1LL << 40;
  co.max_regions = 512, co.max_split_per_step = 64, co.err_smin = 1.0, co.tl_smin = 1.0 / 16;
  co.hmf_smooth = 1, co.init_split_lnm = 4, co.err_scale = 1.5; /* smoothed HMF, lnM split */
  y3o_vegas_opts vo;
  memset(&vo, 0, sizeof(vo));
  vo.eps_rel = 1e-3, vo.eps_abs = 1e-300, vo.max_eval = 1LL << 40;
  vo.calls_per_iter = 4000, vo.itmax = 3, vo.itskip = 1, vo.nbins = 32, vo.alpha = 1.5;
  vo.seed = 0x59334331u;
  vo.hmf_smooth = 1;
  for (int kind = 0; kind < 3; ++kind) {
    y3o_integral ig;
    y3o_make_integral(&p, &s, kind, 20.0, 30.0, 0.2, 0.35, 1.0, &ig);
    y3o_result r;
    y3o_cubature(&p, &t, &ig, &co, &r);
    chk("cubature", r.value);
    y3o_vegas(&p, &t, &ig, &vo, &r);
    chk("vegas", r.value);
  }

  /* the golden integrator on one bin and two radii */
  {
    double const lo[2] = {20.0, 30.0}, zo[2] = {0.2, 0.35}, rr[2] = {0.3, 3.0};
    double counts[1], ds[2];
    y3o_golden_set_orders(2, 3);
    if (y3o_golden(&p, &t, 1, lo, 1, zo, 1, rr, 2, counts, ds)) ++bad; /* 1 x 1 bins */
    chk("golden count", counts[0]);
    chk("golden dsigma", ds[1]);
    y3o_golden_set_orders(4, 6);
  }

  /* the real-data Delta-Sigma integrand and its cubature: small SIG_MAX
   * tables {n_r, n_m, n_z, r, lnM, z, sigma_1, sigma_2, bias} */
  {
    enum { NR = 4, NM = 3, NZ = 3 };
    double des[11 + 3 + NR + NM + NZ + NM * NR + NZ * NR + NM * NZ];
    double const head[11] = {0.2, 3.0, 0.3, 1.0, 0.9, 1e11, 2.5e12, 0.75, 0.2, 0.1, 0.3};
    int k = 0;
    for (int i = 0; i < 11; ++i) des[k++] = head[i];
    des[k++] = NR, des[k++] = NM, des[k++] = NZ;
    for (int i = 0; i < NR; ++i) des[k++] = 0.1 * pow(10.0, i);
    for (int i = 0; i < NM; ++i) des[k++] = 30.0 + 3.0 * i;
    for (int i = 0; i < NZ; ++i) des[k++] = 0.1 + 0.4 * i;
    for (int i = 0; i < NM * NR + NZ * NR + NM * NZ; ++i) des[k++] = 1.0 + 0.01 * i;
    y3o_integral ig;
    memset(&ig, 0, sizeof(ig));
    ig.kind = 3, ig.ndim = 6, ig.zo_lo = 0.3, ig.zo_hi = 0.45, ig.radius = 1.0;
    double const lows[6] = {0.0, 5.0, 0.25, 32.5, 0.2, 0.0};
    double const highs[6] = {600.0, 150.0, 0.5, 33.6, 1.5, 6.283185307179586};
    for (int a = 0; a < 6; ++a) ig.lows[a] = lows[a], ig.highs[a] = highs[a];
    double x[6] = {30.0, 28.0, 0.3, 33.0, 0.5, 1.0};
    chk("integrand_des", y3o_integrand_des(&p, &t, des, &ig, x));
    y3o_result r;
    co.max_regions = 256;
    y3o_cubature_des(&p, &t, des, &ig, &co, &r);
    chk("cubature_des", r.value);
  }

  /* fixed Gauss-Legendre counts with a 3 x 4 selection plane */
  {
    double const lnm[4] = {30.0, 32.0, 34.0, 36.0}, z[3] = {0.1, 0.4, 0.8};
    double plane[12];
    for (int i = 0; i < 12; ++i) plane[i] = 0.1 * i;
    chk("selgl", y3o_selgl_counts(&p, &t, lnm, 4, z, 3, plane, 8, 6, 30.5, 35.0, 0.15, 0.75));
  }

  /* NFW_DSIGMA_MIS: table axes, entries, and a lookup on a 2 x 2 table */
  {
    static double ax[Y3O_NT_NX], bx[Y3O_NT_NM];
    y3o_nfw_mis_axes(ax, bx);
    double T[4];
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) T[j * 2 + i] = y3o_nfw_mis_entry(ax[500 + i], bx[100 + j]);
    chk("nfw table entry", T[3]);
    chk("nfw lookup", y3o_nfw_dsigma_mis_lookup(ax + 500, 2, bx + 100, 2, T, 1.0, 0.3, 33.0,
                                                4.0, 0.3 * 2.77533742639e11));
  }

  y3o_tables_free(&t);
  printf("asan_check: %d non-finite values\n", bad);
  return bad ? 1 : 0;
}
