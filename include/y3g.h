/*
 * y3g.h -- C ABI of the MI355X (gfx950) cluster-observable integration
 * library (liby3g.so).
 *
 * This is the thin layer between the host C++ (the CosmoSIS module template,
 * y3_cluster_cpp_amd/modules/y3g_module.hpp) and the HIP kernels.  Plain pointers and sizes
 * only; every function returns 0 on success and a non-zero y3g_status
 * otherwise, with a message available from y3g_last_error().
 *
 * Reference interfaces each entry point replaces (reference paths relative
 * to the estevesjh/y3_cluster_cpp tree):
 *   y3g_runtime_create      DeviceInitializer, src/utils/CosmoSISSICUDAModule.hh:59-83
 *                           (cudaGetDeviceCount + cudaSetDevice(local_rank))
 *   y3g_tables_build /      the per-sample functor construction of
 *   y3g_tables_upload       set_sample(), e.g. hmf_t.cuh:44-52 (mass_function
 *                           table -> quad::Interp2D on the device),
 *                           dv_do_dz_t.cuh:13-17, and the table copies
 *                           quad::Interp* make inside gpuintegration
 *   y3g_integrate           MultiDimensionalIntegrator<ndim>::integrate,
 *                           src/utils/gpu_integrator.cuh:142-156 ("pagani" ->
 *                           Workspace::integrate, "mcubes" -> mcubes::integrate),
 *                           called once per (volume, grid point) by the serial
 *                           loops of CosmoSISSICUDAModule.hh:282-335; here the
 *                           whole list is one device submission
 *   y3g_result              numint::integration_result fields used at
 *                           CosmoSISSICUDAModule.hh:299-309
 *                           (estimate, errorest, chi_sq, neval, nregions, status)
 *   y3g_des_upload          the DES functors' sample reads in SigmaMiscentY1MortCUDAScalar-
 *                           Integrand::set_sample (sigma_miscent_y1_mort_cuda_scalarintegrand.cu
 *                           :140-150; roffset_t.cuh:9-10, lo_lc_t.cuh:13-18, mor_des_t.cuh
 *                           :264-273, sig_max.cuh:38-57)
 *   y3g_selfn_upload        SelFunction_t's sample reads, src/models/sel_function_t.hh:28-80
 *   y3g_selgl_counts        NumCountsSelGL::set_sample -> SelGLCore::build_weights,
 *                           src/models/n_operator_sel_gl_t.hh:88-150, 184-192
 *                           (all bins of all points in one launch)
 *   y3g_nfw_mis_table_*     NFW_DSIGMA_MIS's table (nfw_dsigma_mis.cuh:29-48,
 *                           62-99: read_vector of the three nfw_off_center
 *                           files into gpu_support::Interp2D)
 *   y3g_eval_functor        the host-side functor calls of the reference's
 *                           Catch2 tests (e.g. test/roffset_t.test.cc:40-46,
 *                           test/int_lc_lt_des_t_test.test.cu for
 *                           INT_LC_LT_DES_t2), here evaluated by the device
 *                           functors
 */
#ifndef Y3G_H
#define Y3G_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  Y3G_OK = 0,
  Y3G_ERR_ARG = 1,      /* invalid argument                              */
  Y3G_ERR_HIP = 2,      /* HIP runtime error                             */
  Y3G_ERR_MEMORY = 3,   /* allocation failed / over the HBM budget       */
  Y3G_ERR_NODEVICE = 4, /* no gfx950 device                              */
  Y3G_ERR_STATE = 5     /* call out of order                             */
} y3g_status;

/* Integrand kinds of the hot path. */
typedef enum {
  Y3G_COUNT3D = 0,  /* (t, zt, lnM), lambda_obs bin analytic (erf)        */
  Y3G_COUNT5D = 1,  /* (lo, zo, t, zt, lnM), z_obs as its standardised
                       photo-z variable over the bin (v in [0, 1])        */
  Y3G_DSIGMA6D = 2, /* (lo, t, zt, lnM, rmis, theta), z_obs bin analytic   */
  /* The reference's real-data (DES) integrands, in the volume's own units,
   * with the point's DES data (y3g_des_upload); FP64 cubature and m-Cubes.
   * The miscentred two integrate lambda_obs in closed form over the
   * descriptor's [lo_lo, lo_hi] (the reference volume's lo range): */
  Y3G_DSIGMA6D_DES = 3, /* (lc, zt, lnM, rmis, theta)      SigmaMiscentY1Mort… */
  Y3G_DSIGMA3D_DES = 4, /* (lo, zt, lnM)                   SigmaCentY1Mort…    */
  Y3G_COUNT3D_DES = 5,  /* (lo, zt, lnM)                   NCCentY1Mort…       */
  Y3G_COUNT5D_DES = 6,  /* (lc, zt, lnM, rmis)             NCMiscentY1Mort…    */
  Y3G_COUNT4D_CARD = 7, /* (ln lo, ln lt, zt, lnM)         numberCountsCard:
                           MOR_DES_LOG_t and INT_LC_LT_DES_t2 tables       */
  /* The Costanzi-2026 P[X] operator integrands P_operator_gpu<W>
   * (p_operator_gpu_t.cuh:118-299), (lt, zt, lnM) with an inner fixed
   * Gauss-Legendre angle integral, grid (zo_low, zo_high, lambda_bin): */
  Y3G_POP_P1 = 8,       /* F = 1                              bSelMargP1GPU  */
  Y3G_POP_I1 = 9,       /* F = b(M, z) xi_NL sigmoid(theta)   bSelMargI1GPU  */
  Y3G_POP_I2 = 10,      /* F = b(M, z) xi_NL                  bSelMargI2GPU  */
  Y3G_POP_J = 11        /* F = b(M, z) xi_NL (1 - sigmoid)    bSelMargJGPU   */
} y3g_kind;

/* Integrators: the names of gpu_integrator.cuh:54-65. */
typedef enum {
  Y3G_PAGANI = 0,   /* Genz-Malik adaptive cubature                       */
  Y3G_MCUBES = 1    /* VEGAS with m-Cubes stratification                  */
} y3g_algorithm;

/* Arithmetic of the integrand. */
typedef enum {
  Y3G_FP32SPLIT = 0, /* FP32 integrand, FP64 sums / errors / grid updates  */
  Y3G_FP64 = 1       /* FP64 throughout                                   */
} y3g_precision;

/* Per-parameter-point physics.  Same field order as the oracle's
 * y3o_params. */
typedef struct {
  double omega_m, h;
  double hmf_amp, hmf_A, hmf_a, hmf_c;
  double mor_lnl0, mor_slope, mor_zslope, mor_sigma_ln;
  double f_prj, prj_scale;
  double f_mis, mis_rlambda;
  double z_source, sigz0;
  double hmf_s, hmf_q;  /* HMF_t amplitude factor s (lnM log10 e - 13.81) + q
                           (hmf_t.cuh:54-60); 0 and 1 for synthetic tables */
} y3g_params;

/* Uniform table grids of one parameter point. */
typedef struct {
  double lnm_lo, lnm_hi;
  int n_lnm;
  double z_lo, z_hi;
  int n_z;
  double da_zmax;
  int n_da;
  double sci_zmax;
  int n_sci;
} y3g_table_spec;

/* One integral.  lows/highs are the box of the integration variables: the
 * MOR variable is integrated as v in [0, 1] across the window of the
 * standardised t = (ln lt - ln lbar) / sigma_ln that lies inside
 * [lnlt_lo, lnlt_hi] and |t| <= 7; for Delta-Sigma the offset radius as its
 * Rayleigh CDF u and the angle on [0, pi]. */
typedef struct {
  int kind;
  int ndim;
  int point;          /* parameter point (table set) index               */
  int algorithm;      /* y3g_algorithm                                   */
  double lo_lo, lo_hi;
  double zo_lo, zo_hi;
  double radius;
  double sigma_mis;
  double rmis_norm;
  double lnlt_lo, lnlt_hi;
  double lows[8], highs[8];
  double eps_rel, eps_abs;
  long long max_eval;
  uint64_t seed;      /* VEGAS Philox key                                */
} y3g_integral;

/* The reference's integration_result. */
typedef struct {
  double value, error, prob;
  long long neval;
  int nregions, status;   /* status 0 = converged, 1 = limit reached,
                             Y3G_STATUS_NO_TABLES: not run (serving: the
                             integral's tables were not ready in time)  */
} y3g_result;
#define Y3G_STATUS_NO_TABLES 3

/* Integrator tuning (defaults from y3g_options_default). */
typedef struct {
  int precision;          /* y3g_precision                               */
  int max_regions;        /* per-integral region capacity (cubature)     */
  int max_split_per_step; /* regions split per step (cubature)           */
  int init_split;         /* initial pieces per axis, 0 = by dimension   */
  int block_threads;      /* threads per workgroup                       */
  int slots;              /* resident workgroups (0 = by device)         */
  long long vegas_calls;  /* VEGAS calls per iteration                   */
  int vegas_itmax, vegas_itskip, vegas_nbins;
  double vegas_alpha;
  int single;             /* 1: one integral per device-wide launch
                             sequence (the plain path)                   */
  double err_c, err_smin; /* region error |R7-R5| * clamp(c r, smin, 1),
                             r = |R7-R5| / |R5-R3|; c = 0 keeps |R7-R5|  */
  int n_bands;            /* tolerance bands (0: every integral gets
                             max_regions): an integral whose eps_rel <=
                             band_eps[k] (first such k, ascending edges)
                             gets a budget of band_regions[k] regions
                             (per-integral call limit 2 x regions x nodes;
                             cubature); from an autotuner                */
  double band_eps[8];
  int band_regions[8];
  double tl_c, tl_smin;   /* two-level estimate (batched cubature): the
                             children of a split region report |R7-R5| x
                             clamp(tl_c d / e, tl_smin, 1), d = |R7(A) +
                             R7(B) - R7(parent)|, e = the parent's |R7-R5|;
                             tl_c = 0 keeps |R7-R5|                     */
  int hmf_smooth;         /* 1 (default): the synthetic integrands
                             (COUNT3D, COUNT5D, DSIGMA6D) read the HMF
                             table through its hat-smoothed cubic spline,
                             whose integral against the integrands' other
                             factors equals the bilinear table's to
                             ~1e-7 relative and on which the cubature
                             converges without the knot-line aliasing of
                             the bilinear form; 0: the bilinear lookup of
                             hmf_t.cuh:54-60 at every node              */
  int init_split_lnm;     /* initial pieces along the lnM axis of the
                             synthetic kinds (default 4; 0: init_split):
                             the MOR and lambda_obs kernels confine the
                             integrands to lnM bands 0.1 - 0.4 wide, which
                             a region spanning half the mass range can
                             hide between its nodes                     */
  double err_scale;       /* safety factor on the cubature's summed error
                             of the integrals that read the smoothed HMF:
                             it stops when err_scale x E + the smoothing
                             bound meets the tolerance and reports that
                             sum (default 1.5).  The bilinear table's
                             kinks used to inflate |R7 - R5| by ~20x;
                             the smooth integrand's estimate is sharper
                             and needs this margin to stay above the
                             true error (0 of 12,288 below it, against 2
                             at 1.0)                                     */
  int ext_regions;        /* budget extension (cubature; 0 = off): an
                             integral whose call limit would cut its
                             next refinement step while its reported
                             error is still above ext_ratio x its
                             tolerance gets the limit doubled, as often
                             as that holds, up to ext_regions regions
                             (<= max_regions); it then refines exactly
                             as with its final limit from the start.
                             The base budget (max_regions or its band)
                             serves the integrals near their tolerance;
                             those left far from it get larger ones     */
  double ext_ratio;
} y3g_options;

typedef struct y3g_runtime y3g_runtime;

const char* y3g_last_error(void);
const char* y3g_version(void);

/* Bind a device (DeviceInitializer semantics: device = local rank). */
int y3g_runtime_create(int device, y3g_runtime** out);
int y3g_runtime_destroy(y3g_runtime* rt);
int y3g_device_info(y3g_runtime* rt, int* n_cu, double* clock_ghz,
                    size_t* hbm_bytes, char* name, int name_len);

void y3g_options_default(y3g_options* o);
void y3g_default_params(y3g_params* p);
void y3g_spec_config0(y3g_table_spec* s);
void y3g_spec_wide(y3g_table_spec* s);

/* Build the tables of n_points parameter points on the device (HMF
 * 200x32, D_A 512, Sigma_crit^-1 128 per point, FP64 and FP32 copies).
 * Replaces any previous point set. */
int y3g_tables_build(y3g_runtime* rt, const y3g_params* params,
                     const y3g_table_spec* spec, int n_points);

/* Upload host tables for point `point` instead of building them (the
 * DataBlock path: mass_function/dndlnmh etc.); arrays sized by the spec. */
int y3g_tables_upload(y3g_runtime* rt, int point, const double* hmf,
                      const double* da, const double* sci);

/* Copy the device tables of one point back (tests). */
int y3g_tables_download(y3g_runtime* rt, int point, double* hmf, double* da,
                        double* sci);

/* Fill an integral descriptor for the synthetic workloads. */
void y3g_make_integral(const y3g_params* p, const y3g_table_spec* s, int kind,
                       double lo_lo, double lo_hi, double zo_lo, double zo_hi,
                       double radius, y3g_integral* out);

/* Fill an integral descriptor from a physical integration volume and a grid
 * point, as a CosmoSIS module reads them: the wall-of-numbers volumes of
 * src/utils/make_cuda_integration_volumes.cuh:48-91 and the grid points of
 * src/utils/make_grid_points.hh:267-284, consumed per integral by
 * CosmoSISSICUDAModule.hh:295-300 (set_grid_point + integrate(volume)).
 * Variables of the volume, in order (lc is the true richness):
 *   Y3G_COUNT3D   (lc, zt, lnm)                  grid (lo_low, lo_high)
 *   Y3G_COUNT5D   (lo, zo, lc, zt, lnm)          no grid
 *   Y3G_DSIGMA6D  (lo, lc, zt, lnm, rmis, theta) grid (zo_low, zo_high, radius)
 * rmis in Mpc/h (the offset pdf is integrated over [rmis_low, rmis_high]),
 * theta in radians inside [0, 2 pi] (with the 1/2pi of the angular average,
 * src/modules/sigma_mort_y1/sigma_miscent_y1_mort_cuda_scalarintegrand.cu:174-175).
 * The real-data kinds take the reference module's volume, (lo, lc, zt, lnm,
 * rmis, theta), (lo, zt, lnm), (lo, zt, lnm), (lo, lc, zt, lnm, rmis) and
 * (ln lo, ln lt, zt, lnm), grid (zo_low, zo_high [, radius]); the
 * miscentred two put its lo range in lo_lo / lo_hi and keep the other axes
 * (lambda_obs in closed form).  The P[X] operator kinds Y3G_POP_* take (lt, zt, lnm) and grid
 * (zo_low, zo_high, lambda_bin), lambda_bin in 0..3 (the DES bins [20, 30),
 * [30, 45), [45, 60), [60, 200] of p_operator_gpu_t.cuh:55-60), and keep the
 * bin in lo_lo and its centre, the integrand's upper lt limit, in lo_hi.
 * Host only; no device is needed. */
int y3g_make_integral_volume(const y3g_params* p, const y3g_table_spec* s,
                             int kind, const double* lows, const double* highs,
                             const double* grid, int n_grid, y3g_integral* out);

/* Integrate n integrals (one device submission, synchronous). */
int y3g_integrate(y3g_runtime* rt, const y3g_integral* igs, size_t n,
                  const y3g_options* opt, y3g_result* out);

/* y3g_integrate without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, validation, tolerance bands and results as
 * y3g_integrate (host descriptors in, host results out); the call's device
 * memory -- the integral and result staging, the work queue, the evaluation
 * counter and the cubature region slab -- is a 64 MiB array that each device
 * holds for the life of the process, cleared where it must start at zero
 * by a fill kernel.  Limits: the staging of n integrals plus one
 * workgroup's region slab (max_regions x 121 B) must fit in the 64 MiB; the
 * cubature keeps as many workgroups resident as the rest of the array holds
 * slabs for (fewer than y3g_integrate at large max_regions).  Rejected with
 * Y3G_ERR_ARG: inputs over those limits, the plain path (single), the
 * real-data kinds.  Every call on one device shares that device's array:
 * calls that overlap on one device (two threads, two runtimes) race. */
int y3g_integrate_noalloc(y3g_runtime* rt, const y3g_integral* igs, size_t n,
                          const y3g_options* opt, y3g_result* out);

/* One batched step as a HIP graph: the device tables of n_points parameter
 * points, then n integrals over them in one persistent cubature launch,
 * captured once over buffers the graph owns (tables, descriptors, results,
 * counters; the region slab is the runtime's) and replayed.  No counterpart
 * in the reference: it is for samplers that repeat a step of a fixed shape,
 * the reference's per-sample loop (CosmoSISSICUDAModule.hh:282-335) with
 * its shape frozen.  Per replay the caller writes the parameter points and
 * the integrals (point = 0 .. n_points - 1; cubature, kinds 0-2) into
 * y3g_graph_params / y3g_graph_integrals, calls y3g_graph_run (synchronous;
 * the host computes the points' constants, the tolerance bands and the
 * longest-first order, then launches the graph) and reads
 * y3g_graph_results.  Results equal y3g_tables_build + y3g_integrate on the
 * same inputs bit for bit; y3g_last_timing reports the replay (table build,
 * copies and launch).  Destroy a graph before its runtime.  A call that reallocates the runtime's region
 * slab (a larger max_regions or more workgroups) invalidates the graph:
 * y3g_graph_run then fails with Y3G_ERR_STATE. */
typedef struct y3g_graph y3g_graph;
int y3g_graph_create(y3g_runtime* rt, const y3g_table_spec* spec, int n_points, size_t n,
                     const y3g_options* opt, y3g_graph** out);
y3g_params* y3g_graph_params(y3g_graph* g);
y3g_integral* y3g_graph_integrals(y3g_graph* g);
const y3g_result* y3g_graph_results(const y3g_graph* g);
int y3g_graph_run(y3g_graph* g);
int y3g_graph_destroy(y3g_graph* g);

/* HBM bytes of the cubature region slab a batched y3g_integrate with these
 * options allocates on this runtime (resident workgroups x max_regions x
 * 121 B); the slab is per resident workgroup, so it does not grow with the
 * number of integrals. */
int y3g_runtime_footprint(y3g_runtime* rt, const y3g_options* opt, size_t* bytes);

/* Device memory as the reference's MemTrackSentry logs it
 * (src/utils/mem_tracking_sentry.hh:37-54, cudaMemGetInfo): the device's free
 * and total bytes, and the bytes this runtime holds (tables, region slab,
 * batch and plain-path buffers, DES, selection and NFW tables). */
int y3g_runtime_memory(y3g_runtime* rt, size_t* free_bytes, size_t* total_bytes,
                       size_t* runtime_bytes);

/* Time of the last y3g_integrate's kernels in milliseconds (HIP events on
 * the library's stream) and the number of integrand evaluations they made. */
int y3g_last_timing(y3g_runtime* rt, double* kernel_ms, long long* nevals);

/* ---- Real-data DES inputs of the Delta-Sigma integrand --------------------
 * The DES data of parameter point `point` for Y3G_DSIGMA6D_DES integrals
 * (the reference's SigmaMiscentY1MortCUDAScalarIntegrand::set_sample,
 * sigma_miscent_y1_mort_cuda_scalarintegrand.cu:140-150):
 *   des = {roffset_tau,                                   ROFFSET_t
 *          LO_LC_alpha, LO_LC_a, LO_LC_b, LO_LC_R_lambda,  LO_LC_t
 *          mor_A, mor_B, mor_alpha, mor_sigma, mor_epsilon, z_mor_pivot,
 *                                                         MOR_DES_t
 *          n_r, n_m, n_z, r[n_r], lnM[n_m], z[n_z],
 *          sigma_1[n_m][n_r], sigma_2[n_z][n_r], bias[n_m][n_z]}   SIG_MAX
 * or, for Y3G_COUNT4D_CARD, MOR_DES_LOG_t's {mor_logMmin, mor_logRatio,
 * mor_alpha, mor_sigma, mor_epsilon, z_mor_pivot} in the MOR_DES places and
 * INT_LC_LT_DES_t2's tables packed as for Y3G_F_INT_LC_LT_DES in place of
 * SIG_MAX.  The counts without SIG_MAX may stop after the 11 parameters.
 * For the P[X] operator kinds (Y3G_POP_*), P_operator_gpu::set_sample's
 * inputs (p_operator_gpu_t.cuh:150-163):
 *   des = {h0, 0, 0, 0, 0,                                cosmological_parameters/h0
 *          log10_Mmin, log10_M1, alpha, sigma_intr, epsilon, z_pivot,
 *                                                         MOR_SHIFTED_POISSON_t
 *          n_chi, n_bm, n_bz, n_xr, n_xz,
 *          chi_z[n_chi], chi_dc[n_chi],                   distances/{z, d_c}
 *          bias_lnM[n_bm], bias_z[n_bz], bias[n_bz][n_bm], haloModel/{lnM, z, bias}
 *          xi_r[n_xr], xi_z[n_xz], xi[n_xz][n_xr]}        xi_nl/{r, z, xi_nl}
 * (the SIG_MAX tables x-fastest, as Interp2D reads the sample's ndarrays).
 * The point's HMF and D_A tables are its y3g_tables_build / _upload ones.
 * Replaces any earlier upload for the point.  y3g_tables_build (every point)
 * and y3g_tables_upload (its point) discard DES data uploaded before them,
 * which belonged to the previous sample: upload it after the tables, and a
 * real-data integral whose point has none fails with Y3G_ERR_STATE. */
int y3g_des_upload(y3g_runtime* rt, int point, const double* des, int n_des);

/* ---- Tabulated miscentred NFW Delta-Sigma -------------------------------
 * The table NFW_DSIGMA_MIS reads (src/models/nfw_dsigma_mis.cuh:29-48,
 * 104-123): T(ln x, ln xmis), the log of the gamma-kernel miscentred
 * Delta-Sigma in units of 2 r_s delta_c rho, in the reference's on-disk
 * layout -- read_vector() of nfw_off_center/table_1000_1e-03_5e+03_
 * gamma_logx.txt (ln x, nx values), ..._gamma_logxmis.txt (ln xmis, nm
 * values) and ..._log_deltasigma_gamma.txt (nm x nx values, x fastest).
 *   y3g_nfw_mis_table_build    builds it on the device (1000 x 250 knots,
 *                              the reference's axes) from the closed-form
 *                              NFW profile, averaged over the offset ring
 *                              and the gamma offset kernel (roffset_t)
 *   y3g_nfw_mis_table_upload   takes the three files' contents instead
 *   y3g_nfw_mis_table_shape / _download  return them (to write the files)
 * Y3G_F_NFW_DSIGMA_MIS evaluates the functor over the runtime's table. */
int y3g_nfw_mis_table_build(y3g_runtime* rt);
int y3g_nfw_mis_table_upload(y3g_runtime* rt, const double* logx, int nx,
                             const double* logxmis, int nm, const double* log_dsigma);
int y3g_nfw_mis_table_shape(y3g_runtime* rt, int* nx, int* nm);
int y3g_nfw_mis_table_download(y3g_runtime* rt, double* logx, double* logxmis,
                               double* log_dsigma);

/* ---- Fixed Gauss-Legendre number counts with a selection function --------
 * The reference's NumCountsSelGL (src/models/n_operator_sel_gl_t.hh:171-201
 * over SelGLCore::build_weights, :88-118): per richness / redshift bin b,
 *   N_b = sum_k w_k sum_q w_q dV/dz(z_q) Omega(z_q) hmf(lnM_k, z_q) S_b(lnM_k, z_q)
 * on n_lnm x n_z Gauss-Legendre nodes over [lnm_lo, lnm_hi] x [zt_lo, zt_hi];
 * Omega(z) is the reference's effective survey area (omega_z_des.hh:15-33),
 * S_b the selection function (sel_function_t.hh:17-60), hmf and dV/dz the
 * point's tables (y3g_tables_build / y3g_tables_upload). */
typedef struct {
  int n_lnm, n_z;            /* Gauss-Legendre nodes (reference default 96, 64) */
  double lnm_lo, lnm_hi;     /* mass window (lnm_low, lnm_high) */
  double zt_lo, zt_hi;       /* redshift window (zt_low, zt_high) */
} y3g_selgl_spec;
#define Y3G_SELGL_MAX_NODES 1024   /* n_lnm, n_z each at most this */
/* The selection function of parameter point `point`: sel_function/lnM
 * [n_lnm], z [n_z] (both increasing) and S_stack [n_bins][n_z][n_lnm]
 * (sel_function_t.hh:44-60).  Replaces any earlier upload for the point. */
int y3g_selfn_upload(y3g_runtime* rt, int point, const double* lnm, int n_lnm,
                     const double* z, int n_z, const double* s_stack, int n_bins);
/* N[point * max_bins + b] for points [0, n_points) and each point's bins
 * b < its n_bins (entries beyond a point's bins are set to 0).  Every point
 * needs tables and a selection function.  FP64 throughout. */
int y3g_selgl_counts(y3g_runtime* rt, const y3g_selgl_spec* spec, int n_points,
                     int max_bins, double* out);

/* Evaluate a device functor on n inputs (tests of the model functors).
 * which: see y3g_functor; in[n * nin], out[n]. */
typedef enum {
  Y3G_F_ROFFSET_GAMMA = 0, /* (x; tau)                          roffset_t   */
  Y3G_F_LO_LC = 1,         /* (lo, lc, R; alpha, a, b, R_lambda) lo_lc_t    */
  Y3G_F_ZO_ZT_DES = 2,     /* (zo, zt)                          zo_zt_des_t */
  Y3G_F_INT_ZO_ZT_DES = 3, /* (zomin, zomax, zt)          int_zo_zt_des_t   */
  Y3G_F_SIGMA_PHOTOZ = 4,  /* (zt)                        sigma_photoz_des  */
  Y3G_F_DV_DO_DZ = 5,      /* (zt) of point 0                 dv_do_dz_t    */
  Y3G_F_HMF = 6,           /* (lnM, zt) of point 0                hmf_t     */
  Y3G_F_MOR = 7,           /* (lt, lnM, zt) of point 0    lognormal MOR     */
  Y3G_F_LO_LT_PDF = 8,     /* (lo, lt) of point 0   Gauss + EMG tail        */
  Y3G_F_LO_LT_BIN = 9,     /* (lo_lo, lo_hi, lt) of point 0                 */
  Y3G_F_NFW_DSIGMA = 10,   /* (R, lnM, rho, conc)  centred NFW Delta-Sigma  */
  Y3G_F_DSIGMA_OFFSET = 11,/* (R, rmis, theta, lnM, rho, conc) tangential    */
  Y3G_F_SCI = 12,          /* (zt) of point 0   Sigma_crit^-1              */
  Y3G_F_INTEGRAND = 13,    /* (y[ndim]) mapped integrand of integral `aux` */
  Y3G_F_INTEGRAND_PAIRED = 14,/* the same through pair_factors + integrand_pair,
                                 the cubature's factorised form       */
  Y3G_F_INTEGRAND_NODE = 15,  /* the FP32 cubature's node form: pair factors
                                 at (zt, lnM), the other coordinates
                                 rounded to float (FP64: as 14)        */
  Y3G_F_INT_LC_LT_DES = 16,   /* (lc, lt, zt); par = {n_lt, n_zt, lt knots,
                                 zt knots, 4 tables of n_lt x n_zt, lt
                                 fastest}           INT_LC_LT_DES_t2, FP64 */
  Y3G_F_MOR_DES = 17,         /* (lt, lnM, zt); par = {A, B, alpha, sigma_intr,
                                 epsilon, z_pivot}  MOR_DES_t, FP64        */
  Y3G_F_EMG_DES_PDF = 18,     /* (lob, ltr, z); par = {n, z[n], a_mu, b_mu,
                                 a_sig, b_sig, a_tau, b_tau, a_fprj, b_fprj
                                 (n each)}          EMG_DES_t(), FP64      */
  Y3G_F_EMG_DES_CDF = 19,     /* the same arguments  EMG_DES_t::cdf, FP64  */
  Y3G_F_SIG_MAX = 20,         /* (r, lnM, zt); par = {n_r, n_m, n_z, r, lnM, z,
                                 sigma_1 [n_m][n_r], sigma_2 [n_z][n_r],
                                 bias [n_m][n_z]}   SIG_MAX, FP64          */
  Y3G_F_NFW_DSIGMA_MIS = 22,  /* (r, rmis, lnM); par = {c, rho}: NFW_DSIGMA_MIS
                                 over the runtime's table (y3g_nfw_mis_table_*),
                                 FP64 */
  Y3G_F_SCI_AVERAGE = 24,     /* (zl); par = {n, zlense[n], sci_average[n]}:
                                 the sample's average_sigma_crit_inv table,
                                 Interp1D clamp (lensing_weights.hh:55-60, 308) */
  Y3G_F_SCI_RZ = 25,          /* (r, zt); par = {n_r, n_z, r[n_r], z[n_z],
                                 sigma_crit_inv[n_z][n_r]}: Interp2D clamp(r, z)
                                 (gamma_1h_nfw.cuh:81-88, 120) */
  Y3G_F_LC_LT = 23,           /* (lc, lt, zt)   LC_LT_t over its fitted tables
                                 (lc_lt_t.cuh:36-68, lc_lt_t.cc:10-214) */
  Y3G_F_MOR_DES_LOG = 21      /* (lt, lnM, zt); par = {log10 Mmin, log10 ratio,
                                 alpha, sigma_intr, epsilon, z_pivot}
                                                    MOR_DES_LOG_t, FP64    */
} y3g_functor;
int y3g_eval_functor(y3g_runtime* rt, int which, int precision,
                     const double* in, int nin, size_t n, const double* par,
                     int npar, const y3g_integral* aux, double* out);

/* ------------------------------------------------------------------ serving
 * Many parameter points in flight at once (samplers with many walkers or
 * chains, several likelihood processes sharing one GPU).  The reference
 * evaluates one sample per execute(), each a serial loop over its
 * integrals (src/utils/CosmoSISSICUDAModule.hh:252-271, 282-335); here a
 * request is one parameter point plus its integrals, and the server runs
 * ONE persistent cubature launch that takes integrals from a ring in
 * pinned host memory as requests arrive (continuous batching: a request
 * joins the launch that is already running, and its integrals start as
 * soon as a workgroup is free).
 *
 * Tables: each parameter point's tables live in a table slot on the device.
 * Slots are kept resident under an LRU, so a repeated cosmology reuses its
 * tables; a new one takes a free or least-recently-used idle slot, and its
 * table build runs inside the same launch, ahead of its integrals.
 * Admission: the region slabs of the resident workgroups and table_slots
 * table sets must fit hbm_budget_bytes (fewer slots are kept if they do not);
 * a request is admitted, in arrival order, when a slot and ring_integrals
 * room are free.  A launch with nothing to do ends after idle_ms and the
 * next request starts another.  Results of (parameter point, integral)
 * pairs are kept in an LRU cache of cache_entries entries.  The integrator
 * is the cubature (Y3G_PAGANI); results are written into the caller's array
 * before y3g_server_wait returns, and that array must stay alive until
 * then. */
typedef struct y3g_server y3g_server;

typedef struct {
  int ring_integrals;         /* integrals (and table builds) in flight     */
  int table_slots;            /* resident per-cosmology table sets          */
  size_t hbm_budget_bytes;    /* bound on region slabs + table sets         */
  long long cache_entries;    /* LRU result cache size, 0 = off             */
  double idle_ms;             /* an idle launch ends after this long        */
  double table_wait_ms;       /* bound on an integral's wait for its tables */
  y3g_options options;        /* integrator options (precision fp32split or
                                 fp64; block_threads is 512)               */
  y3g_table_spec spec;        /* table grids of every parameter point      */
} y3g_server_options;

typedef struct {
  long long requests_submitted, requests_done;
  long long integrals_computed, cache_hits, evaluations, queued;
  long long launches;         /* persistent launches started               */
  long long table_builds, table_reuses;
  long long slab_bytes, table_bytes;  /* device memory of the session      */
  int table_slots;            /* table slots admitted by the HBM budget    */
  int workgroups;             /* resident workgroups of the launch         */
} y3g_server_stats;

const char* y3g_server_last_error(void);
void y3g_server_options_default(y3g_server_options* o);
int y3g_server_create(int device, const y3g_server_options* opt, y3g_server** out);
/* Queue one request; integral `point` fields are ignored (the server
 * assigns table slots).  Cache hits are written to out at once.  Only
 * Y3G_PAGANI integrals of the synthetic kinds (0-2) are served; a request
 * larger than the ring is refused. */
int y3g_server_submit(y3g_server* s, const y3g_params* params,
                      const y3g_integral* igs, size_t n, y3g_result* out,
                      unsigned long long* ticket);
/* Wait for a request (timeout_s < 0: no limit); latency_s = completion
 * minus submission on the server's steady clock.  A ticket is waited once. */
int y3g_server_wait(y3g_server* s, unsigned long long ticket, double timeout_s,
                    double* latency_s);
int y3g_server_stats_get(y3g_server* s, y3g_server_stats* st);
/* Finish the queued requests, end the launch, free everything. */
int y3g_server_destroy(y3g_server* s);

#ifdef __cplusplus
}
#endif
#endif
