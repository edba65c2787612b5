/*
 * y3o.h -- FP64 CPU oracle for the cluster-observable integration path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product path (y3_cluster_cpp_amd, the HIP library and
 * the CosmoSIS modules) never links or calls it.
 *
 * It restates, in plain C99 and double precision:
 *   - the reference model functors (reference form, for the reference's own
 *     known-answer fixtures) and their synthetic-workload parameterisations;
 *   - the three integrands of the hot path (3-D count, 5-D count, 6-D
 *     miscentred Delta-Sigma);
 *   - a structure-exploiting nested Gauss-Legendre / Gauss-Kronrod "golden"
 *     integrator (truth to ~1e-9);
 *   - CPU restatements of the two integrators the GPU library implements
 *     (Genz-Malik adaptive cubature, VEGAS with m-Cubes stratification), with
 *     the same region / grid algorithms and the same Philox4x32-10 streams.
 *
 * Parity pins (which check pins which function) are listed in the header
 * comment of y3o.c next to each function.
 */
#ifndef Y3O_H
#define Y3O_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- parameters ------------------------------------------- */

/* Per-parameter-point physics (one "cosmology" of the sampler).  Field order
 * is shared with include/y3g.h (y3g_params) so one marshalling serves both. */
typedef struct {
  double omega_m;        /* flat LCDM, Omega_Lambda = 1 - omega_m          */
  double h;              /* H0 / 100                                        */
  double hmf_amp;        /* sigma(M,z) amplitude (0.8)                      */
  double hmf_A;          /* 0.2                                             */
  double hmf_a;          /* 1.5                                             */
  double hmf_c;          /* 1.2                                             */
  double mor_lnl0;       /* ln lambda-bar at pivot (3.2)                    */
  double mor_slope;      /* d ln lambda / d ln M (0.9)                      */
  double mor_zslope;     /* d ln lambda / d ln(1+z) (0.5)                   */
  double mor_sigma_ln;   /* lognormal scatter (0.25)                        */
  double f_prj;          /* projection-tail fraction (0.2; 0 for 3-D)       */
  double prj_scale;      /* exponential excess scale / lambda_true (0.1)    */
  double f_mis;          /* miscentred fraction (0.2)                       */
  double mis_rlambda;    /* sigma_Rayleigh / R_lambda (0.2)                 */
  double z_source;       /* source plane for Sigma_crit^-1 (1.2)            */
  double sigz0;          /* photo-z scatter sigma_z = sigz0 (1+z) (0.01)    */
  double hmf_s, hmf_q;   /* HMF_t amplitude factor s (lnM log10 e - 13.81)
                            + q (hmf_t.cuh:54-60); 0 and 1 for the
                            synthetic tables                               */
} y3o_params;

/* Table geometry (uniform grids). */
typedef struct {
  double lnm_lo, lnm_hi;  int n_lnm;   /* HMF lnM axis, 200              */
  double z_lo, z_hi;      int n_z;     /* HMF z axis, 32                 */
  double da_zmax;         int n_da;    /* D_A(z) on [0, da_zmax], 512    */
  double sci_zmax;        int n_sci;   /* Sigma_crit^-1(z_l), 128        */
} y3o_table_spec;

/* Tables of one parameter point, FP64.  hmf is stored x-fastest:
 * hmf[iz * n_lnm + im]  (ndarray of shape (n_z, n_lnM)). */
typedef struct {
  y3o_table_spec spec;
  double *lnm, *z, *hmf;     /* n_lnm, n_z, n_z*n_lnm */
  double *da_z, *da;         /* n_da, Mpc             */
  double *sci_z, *sci;       /* n_sci                 */
  /* Cubic-spline B-coefficients of the HMF table, (n_z + 4) x (n_lnm + 4),
   * index (l + 2) (n_lnm + 4) + (k + 2) for l in [-2, n_z + 1], k in
   * [-2, n_lnm + 1] (y3o_hmf_coefs), and whether the integrands evaluate
   * the HMF through them (y3o_hmf_smooth) instead of bilinearly. */
  double *hmfc;
  int smooth_hmf;
} y3o_tables;

/* Integrand kinds. */
enum { Y3O_COUNT3D = 0, Y3O_COUNT5D = 1, Y3O_DSIGMA6D = 2 };

/* One integral of the hot path. */
typedef struct {
  int kind;
  int ndim;
  double lo_lo, lo_hi;     /* lambda_obs bin                                */
  double zo_lo, zo_hi;     /* z_obs bin                                     */
  double radius;           /* R [Mpc/h] for Delta-Sigma                     */
  double sigma_mis;        /* Rayleigh sigma [Mpc/h] for Delta-Sigma        */
  double rmis_norm;        /* 1 / (1 - exp(-rmax^2 / 2 sigma^2))            */
  double lt_lo, lt_hi;     /* lambda_true range [1, 300]                    */
  double lows[8], highs[8];  /* box of the integration variables y        */
} y3o_integral;

/* Integration result: the fields of the reference's integration_result. */
typedef struct {
  double value, error, prob;
  long long neval;
  int nregions, status;
} y3o_result;

/* ---------------- primitives and reference-form functors --------------- */
double y3o_gaussian(double x, double mu, double sigma);
double y3o_roffset_gamma(double x, double tau);
double y3o_lo_lc(double lo, double lc, double R_mis, double alpha, double a,
                 double b, double R_lambda);
double y3o_lo_lc_bin(double lo_lo, double lo_hi, double lc, double R_mis, double alpha,
                     double a, double b, double R_lambda);
double y3o_sigma_photoz_des(double zt);
double y3o_zo_zt_des(double zo, double zt);
double y3o_int_zo_zt_des(double zomin, double zomax, double zt);
double y3o_ez(double z, double om, double ol, double ok);
int    y3o_interp1d(const double* xs, const double* ys, int n, double x,
                    double* out);
double y3o_interp1d_clamp(const double* xs, const double* ys, int n, double x);
int    y3o_interp2d(const double* xs, int nx, const double* ys, int ny,
                    const double* zs, double x, double y, double* out);
double y3o_interp2d_clamp(const double* xs, int nx, const double* ys, int ny,
                          const double* zs, double x, double y);
double y3o_int_lc_lt_des(const double* par, double lc, double lt, double zt);
void   y3o_mor_des_sigma_skew(double sigma_intr, double lam, double* sigma, double* skw);
double y3o_mor_des(const double* q, double lt, double lnM, double zt);
double y3o_mor_des_log(const double* q, double lt, double lnM, double zt);
void   y3o_emg_des_params(const double* par, double ltr, double z, double* out);
double y3o_emg_des_cdf(const double* par, double lob, double ltr, double z);
double y3o_emg_des_pdf(const double* par, double lob, double ltr, double z);
double y3o_sig_max(const double* par, double r, double lnM, double zt);
double y3o_dv_do_dz(double zt, const double* da_z, const double* da, int n,
                    double om, double ol, double ok, double h);
double y3o_lc_lt_eq15(double lc, double lt, double mu, double sigma,
                      double tau, double fmsk, double fprj);

/* ---------------- synthetic-workload functors --------------------------- */
double y3o_hmf_formula(const y3o_params* p, double lnM, double z);
double y3o_mor_lognormal(const y3o_params* p, double lt, double lnM, double zt);
double y3o_lo_lt_pdf(const y3o_params* p, double lo, double lt);
double y3o_lo_lt_bin(const y3o_params* p, double lo_lo, double lo_hi, double lt);
double y3o_emg_pdf(double x, double mu, double sigma, double tau);
double y3o_emg_cdf(double x, double mu, double sigma, double tau);
double y3o_zo_pdf(const y3o_params* p, double zo, double zt);
double y3o_zo_bin(const y3o_params* p, double zo_lo, double zo_hi, double zt);
double y3o_comoving_distance(double om, double h, double z); /* Mpc */
double y3o_nfw_conc(double lnM);
double y3o_nfw_g(double x);  /* Delta-Sigma / (rho_s r_s) */
double y3o_nfw_sigma_unit(double x); /* Sigma / (rho_s r_s) */
double y3o_nfw_dsigma(double R, double lnM, double rho, double conc);
double y3o_dsigma_offset(double R, double rmis, double theta, double lnM,
                         double rho, double conc);
double y3o_dsigma_mis_ring(double R, double rmis, double lnM, double rho,
                           double conc, double epsrel);
double y3o_dsigma_mis_gamma(double R, double tau, double lnM, double rho,
                            double conc, double epsrel);
double y3o_dsigma_mis_rayleigh(double R, double sigma, double rmax,
                               double lnM, double rho, double conc,
                               double epsrel);

/* ---------------- tables ------------------------------------------------ */
void y3o_default_params(y3o_params* p);
void y3o_spec_config0(y3o_table_spec* s);
void y3o_spec_wide(y3o_table_spec* s);
int  y3o_tables_build(const y3o_params* p, const y3o_table_spec* s,
                      y3o_tables* t);
void y3o_tables_free(y3o_tables* t);
double y3o_hmf(const y3o_params* p, const y3o_tables* t, double lnM, double zt);
/* The hat-smoothed HMF: the cubic spline through the table convolved with
 * the bilinear kernel of one cell per axis, as a tensor quintic B-spline
 * over the spline's coefficients (y3o_hmf_coefs).  Integrated against any
 * factor smooth on the scale of a table cell it gives the integral of the
 * bilinear table to a Poisson-summation remainder; see y3o_models.c. */
double y3o_hmf_smooth(const y3o_params* p, const y3o_tables* t, double lnM, double zt);
void   y3o_spline_coefs(const double* y, int n, long stride, double* c, long cstride);
void   y3o_hmf_coefs(const double* hmf, int n_lnm, int n_z, double* c);
void   y3o_b5_weights(double f, double* w);
/* Bound on the smoothing identity's remainder relative to the integral
 * (1e-6, against at most 1.2e-7 measured), added by the cubature and VEGAS
 * to the error they test and report for the synthetic kinds when these read
 * the smoothed HMF; 0 otherwise (csrc/y3g_device.hpp smooth_bound). */
double y3o_smooth_bound(int kind, int smooth);
double y3o_dvdz(const y3o_params* p, const y3o_tables* t, double zt);
double y3o_sci(const y3o_tables* t, double zt);

/* ---------------- integrands ------------------------------------------ */
double y3o_integrand(const y3o_params* p, const y3o_tables* t,
                     const y3o_integral* ig, const double* x);
double y3o_integrand_mapped(const y3o_params* p, const y3o_tables* t,
                            const y3o_integral* ig, const double* y);
void   y3o_integrand_mapped_batch(const y3o_params* p, const y3o_tables* t,
                                  const y3o_integral* ig, const double* y,
                                  long n, double* out);
void   y3o_make_integral(const y3o_params* p, const y3o_table_spec* s,
                         int kind, double lo_lo, double lo_hi, double zo_lo,
                         double zo_hi, double radius, y3o_integral* ig);

/* ---------------- quadrature ------------------------------------------ */
double y3o_gk21_adapt(double (*f)(double, void*), void* ctx, double a,
                      double b, double epsrel, double epsabs, double* err);
void   y3o_gl_rule(int n, double* x, double* w);
const double* y3o_gk21_nodes(void);
const double* y3o_gk21_wk(void);
const double* y3o_gk21_wg(void);

/* Golden values: counts[n_lo*n_zo] (config 0 when kind == COUNT3D and
 * n_zo == 0) and, when n_r > 0, dsig[n_lo*n_zo*n_r] (bin-major, radius
 * fastest).  Returns 0 on success. */
int y3o_golden(const y3o_params* p, const y3o_tables* t, int count_kind,
               const double* lo_edges, int n_lo, const double* zo_edges,
               int n_zo, const double* radii, int n_r, double* counts,
               double* dsig);

void y3o_golden_set_orders(int ngz, int ngm);

/* Genz-Malik adaptive cubature, the CPU restatement of the device
 * algorithm (csrc/y3g_kernels.hip k_cubature). */
typedef struct {
  double eps_rel, eps_abs;
  long long max_eval;
  int max_regions;
  int max_split_per_step;
  int init_split;           /* pieces per axis of the initial partition */
  double err_c, err_smin;   /* region error E1 clamp(c E1/E2, smin, 1)     */
  double tl_c, tl_smin;     /* two-level child scale clamp(c d/e, smin, 1) */
  int hmf_smooth;           /* integrands read the hat-smoothed HMF        */
  int init_split_lnm;       /* initial pieces along the synthetic kinds'
                               lnM axis (0: init_split)                    */
  double err_scale;         /* safety factor on E (tested and reported)    */
  int ext_regions;          /* budget extension: a call limit about to cut */
  double ext_ratio;         /* a step with the error above ext_ratio x the
                               tolerance is doubled, up to 2 ext_regions
                               nodes (0: off)                              */
} y3o_cub_opts;
void y3o_cubature(const y3o_params* p, const y3o_tables* t,
                  const y3o_integral* ig, const y3o_cub_opts* o,
                  y3o_result* r);
/* The real-data Delta-Sigma integrand (Y3G_DSIGMA6D_DES) and its cubature
 * (y3o_models.c). */
double y3o_integrand_des(const y3o_params* p, const y3o_tables* t, const double* des,
                         const y3o_integral* ig, const double* x);
void   y3o_cubature_des(const y3o_params* p, const y3o_tables* t, const double* des,
                        const y3o_integral* ig, const y3o_cub_opts* o, y3o_result* r);
/* Generic: integrate f over [lows, highs]. */
void y3o_cubature_fn(double (*f)(const double*, void*), void* ctx, int ndim,
                     const double* lows, const double* highs,
                     const y3o_cub_opts* o, y3o_result* r);
/* Genz-Malik rule of degree 7 / 5 on one region; returns R7, sets *r5 and
 * the fourth differences fourth[ndim]. */
double y3o_gm_rule(double (*f)(const double*, void*), void* ctx, int ndim,
                   const double* center, const double* halfw, double* r5,
                   double* fourth);
int y3o_gm_num_nodes(int ndim);
double y3o_gm_rule3(double (*f)(const double*, void*), void* ctx, int ndim,
                    const double* center, const double* halfw, double* r5,
                    double* r3, double* fourth);

/* VEGAS / m-Cubes. */
typedef struct {
  double eps_rel, eps_abs;
  long long max_eval;       /* maxcalls */
  long long calls_per_iter;
  int itmax, itskip, nbins;
  double alpha;
  uint64_t seed;
  int hmf_smooth;           /* integrands read the hat-smoothed HMF        */
} y3o_vegas_opts;
void y3o_vegas(const y3o_params* p, const y3o_tables* t,
               const y3o_integral* ig, const y3o_vegas_opts* o, y3o_result* r);
void y3o_vegas_fn(double (*f)(const double*, void*), void* ctx, int ndim,
                  const double* lows, const double* highs,
                  const y3o_vegas_opts* o, y3o_result* r);

/* Philox4x32-10 (Salmon et al. 2011). */
void y3o_philox(uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Fixed Gauss-Legendre number counts with a selection function
 * (n_operator_sel_gl_t.hh NumCountsSelGL; y3o_selgl.c). */
double y3o_omega_z_des(double z);
double y3o_sel_function(const double* lnm, int nm, const double* z, int nz,
                        const double* plane, double lnM, double zt);
void   y3o_gl_nodes_ab(int n, double a, double b, double* x, double* w);
double y3o_selgl_counts(const y3o_params* p, const y3o_tables* t,
                        const double* lnm, int nm, const double* z, int nz,
                        const double* plane, int n_lnm, int n_z, double lnm_lo,
                        double lnm_hi, double zt_lo, double zt_hi);

/* Test integrands of the reference's integrator known-answer tests. */
double y3o_test_fn(int which, const double* x, int ndim);

/* The Costanzi-2026 P[X] operator integrands (y3o_pop.c); kinds 8..11 of
 * y3o_integrand_des / y3o_cubature_des. */
double y3o_sigma_photoz_table(double z);
double y3o_mor_shifted_poisson(const double* q, double ltr, double lnM, double z);
double y3o_pop_area_overlap(double th, double lob, double lt);
double y3o_pop_theta_excl(double chi_z, double chi_o, double R_excl);
double y3o_pop_theta_term(int kind, double th, double theta_lob, double theta_lt,
                          double chi_z, double chi_o, double zob, const double* xr,
                          int nxr, const double* xz, int nxz, const double* xt);
double y3o_integrand_pop(const y3o_params* p, const y3o_tables* t, const double* des,
                         const y3o_integral* ig, const double* x);

/* LC_LT_t over the reference's fitted tables (y3o_models.c). */
double y3o_lc_lt(double lc, double lt, double zt);
void   y3o_lc_lt_params(double lt, double zt, double* out5);

/* Tabulated miscentred NFW Delta-Sigma, NFW_DSIGMA_MIS (y3o_nfwtab.c). */
#define Y3O_NT_NX 1000
#define Y3O_NT_NM 250
#define Y3O_NT_MAXPAN 128
int    y3o_nfw_ring_levels(double x, double s);
double y3o_nfw_ring(double x, double s);
int    y3o_nfw_mis_panels(double x, double m, double* lo, double* hi);
double y3o_nfw_mis_unit(double x, double m);
double y3o_nfw_mis_unit_abs(double x, double m);
void   y3o_nfw_mis_axes(double* logx, double* logxmis);
double y3o_nfw_mis_entry(double logx, double logxmis);
double y3o_nfw_dsigma_mis_lookup(const double* logx, int nx, const double* logxmis, int nm,
                                 const double* T, double r, double rmis, double lnM,
                                 double c, double rho);

#ifdef __cplusplus
}
#endif
#endif
