/*
 * y3o_batch.c -- OpenMP batch drivers over the oracle's CPU integrators.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the CPU baseline timed by bench.py
 * (the FP64 restatement integrated by CPU Genz-Malik cubature / VEGAS,
 * one integral per thread, like the reference CPU path's one integral per
 * call of CosmoSISScalarIntegrationModule.hh:162-190) and the golden-value
 * driver used by tests/golden/make_golden.py.
 */
#include "y3o.h"

#include <omp.h>
#include <stdlib.h>

/* Integrate n integrals; integral i uses params[point[i]] and
 * tables[point[i]], eps_rel[i] and (cubature) max_eval[i] when those arrays
 * are given.  algo 0 = cubature, 1 = VEGAS.  Returns the thread count used. */
int
y3o_batch_integrate(const y3o_params* params, const y3o_table_spec* specs,
                    int n_points, const y3o_integral* igs, const int* point,
                    const double* eps_rel, const long long* max_eval, int n, int algo,
                    const y3o_cub_opts* copt, const y3o_vegas_opts* vopt,
                    int nthreads, y3o_result* out)
{
  y3o_tables* t = (y3o_tables*)calloc(n_points, sizeof(y3o_tables));
  if (nthreads > 0) omp_set_num_threads(nthreads);
  int used = 1;
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < n_points; ++k) y3o_tables_build(&params[k], &specs[k], &t[k]);
#pragma omp parallel
  {
#pragma omp single
    used = omp_get_num_threads();
#pragma omp for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) {
      int const k = point[i];
      if (algo == 0) {
        y3o_cub_opts o = *copt;
        if (eps_rel) o.eps_rel = eps_rel[i];
        if (max_eval) o.max_eval = max_eval[i];
        y3o_cubature(&params[k], &t[k], &igs[i], &o, &out[i]);
      } else {
        y3o_vegas_opts o = *vopt;
        if (eps_rel) o.eps_rel = eps_rel[i];
        o.seed = vopt->seed + (uint64_t)i;
        y3o_vegas(&params[k], &t[k], &igs[i], &o, &out[i]);
      }
    }
  }
  for (int k = 0; k < n_points; ++k) y3o_tables_free(&t[k]);
  free(t);
  return used;
}

/* Golden values for n_points parameter points, each with the full
 * lambda_obs x z_obs x R grid.  Output layout per point: counts
 * [n_lo * n_zo] then dsig [n_lo * n_zo * n_r]. */
int
y3o_batch_golden(const y3o_params* params, const y3o_table_spec* spec,
                 int n_points, int count_kind, const double* lo_edges,
                 int n_lo, const double* zo_edges, int n_zo,
                 const double* radii, int n_r, int nthreads, double* out)
{
  if (nthreads > 0) omp_set_num_threads(nthreads);
  int const nzb = n_zo > 0 ? n_zo : 1;
  int const stride = n_lo * nzb * (1 + n_r);
  int rc = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : rc)
  for (int k = 0; k < n_points; ++k) {
    y3o_tables t;
    y3o_tables_build(&params[k], spec, &t);
    double* o = out + (size_t)k * stride;
    rc |= y3o_golden(&params[k], &t, count_kind, lo_edges, n_lo, zo_edges,
                     n_zo, radii, n_r, o, n_r > 0 ? o + n_lo * nzb : 0);
    y3o_tables_free(&t);
  }
  return rc;
}
