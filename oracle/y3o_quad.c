/*
 * y3o_quad.c -- FP64 quadrature for the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Never linked by the product.
 *
 *   y3o_gk21_adapt   adaptive Gauss-Kronrod 21 (QUADPACK QK21 constants);
 *                    pinned by polynomial exactness (degree 31 / 19) and the
 *                    1-D closed forms of integration.test.cc (sf1 = 0.5).
 *   y3o_gl_rule      Gauss-Legendre by Newton on P_n; pinned by exactness of
 *                    degree 2n-1 monomials.
 *   y3o_golden       nested golden integrator (see comment at the function);
 *                    pinned against brute-force 5-D / 6-D cubature of the
 *                    full integrand and against the chebyshev-free direct
 *                    inner integral at sample points.
 *   y3o_gm_rule      Genz & Malik (1980) degree-7 rule with embedded degree 5;
 *                    pinned by exactness on all monomials of degree <= 7 / 5.
 *   y3o_cubature*    adaptive cubature (restatement of the device algorithm);
 *                    pinned by integration.test.cc (sf2 = 1, sf2_a_b = -1276,
 *                    sf6 = 4.14199) and ExampleScalarIntegrand.test.cc
 *                    (0.345833, 346.688).
 *   y3o_vegas*       VEGAS (Lepage 1978) with m-Cubes stratification;
 *                    pinned by the same closed forms.
 *   y3o_philox       Philox4x32-10 (Salmon et al. 2011); pinned by the
 *                    Random123 known-answer vectors.
 */
#include "y3o.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define Y3O_PI 3.14159265358979323846
#define Y3O_RHOC 2.77533742639e+11

/* ---------------- Gauss-Kronrod 21 ------------------------------------- */
static const double xgk[11] = {
  0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
  0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
  0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
  0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
  0.294392862701460198131126603103866, 0.148874338981631210884826001129720,
  0.000000000000000000000000000000000};
static const double wgk[11] = {
  0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
  0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
  0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
  0.123491976262065851077600718054245, 0.134709217311473325928054001771707,
  0.142775938577060080797094273138717, 0.147739104901338491374841515972068,
  0.149445554002916905664936468389821};
static const double wg[5] = {
  0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
  0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
  0.295524224714752870173892994651338};

const double* y3o_gk21_nodes(void) { return xgk; }
const double* y3o_gk21_wk(void) { return wgk; }
const double* y3o_gk21_wg(void) { return wg; }

static double
gk21(double (*f)(double, void*), void* ctx, double a, double b, double* err)
{
  double const c = 0.5 * (a + b), h = 0.5 * (b - a);
  double const fc = f(c, ctx);
  double rk = fc * wgk[10], rg = 0.0;
  for (int j = 0; j < 10; ++j) {
    double const dx = h * xgk[j];
    double const f1 = f(c - dx, ctx), f2 = f(c + dx, ctx);
    rk += wgk[j] * (f1 + f2);
    if (j & 1) rg += wg[j / 2] * (f1 + f2);
  }
  *err = fabs((rk - rg) * h);
  return rk * h;
}

typedef struct {
  double a, b, v, e;
} seg_t;

/* Globally adaptive bisection (QUADPACK QAG strategy): split the interval
 * with the largest error until sum(err) <= max(epsabs, epsrel |I|). */
double
y3o_gk21_adapt(double (*f)(double, void*), void* ctx, double a, double b,
               double epsrel, double epsabs, double* err)
{
  int cap = 4096, n = 1;
  seg_t* s = (seg_t*)malloc(sizeof(seg_t) * cap);
  s[0].a = a;
  s[0].b = b;
  s[0].v = gk21(f, ctx, a, b, &s[0].e);
  double I = s[0].v, E = s[0].e;
  while (E > fmax(epsabs, epsrel * fabs(I)) && n < cap) {
    int w = 0;
    for (int i = 1; i < n; ++i)
      if (s[i].e > s[w].e) w = i;
    double const m = 0.5 * (s[w].a + s[w].b);
    if (!(m > s[w].a && m < s[w].b)) break;
    seg_t l = {s[w].a, m, 0, 0}, r = {m, s[w].b, 0, 0};
    l.v = gk21(f, ctx, l.a, l.b, &l.e);
    r.v = gk21(f, ctx, r.a, r.b, &r.e);
    s[w] = l;
    s[n++] = r;
    I = 0.0;
    E = 0.0;
    for (int i = 0; i < n; ++i) I += s[i].v, E += s[i].e;
  }
  free(s);
  if (err) *err = E;
  return I;
}

/* Gauss-Legendre nodes/weights on [-1, 1] by Newton iteration on P_n. */
void
y3o_gl_rule(int n, double* x, double* w)
{
  for (int i = 0; i < n; ++i) {
    double z = cos(Y3O_PI * (i + 0.75) / (n + 0.5)), pp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        double const p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      double const z1 = z;
      z = z1 - p1 / pp;
      if (fabs(z - z1) < 1e-16) break;
    }
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
}

/* ---------------- miscentred rings -------------------------------------- */
typedef struct {
  double R, rmis, lnM, rho, conc;
  double sig, rmax, tau;
  int kernel; /* 0 rayleigh, 1 gamma */
  double epsrel;
} ring_ctx;

static double
ring_theta(double th, void* v)
{
  ring_ctx* c = (ring_ctx*)v;
  return y3o_dsigma_offset(c->R, c->rmis, th, c->lnM, c->rho, c->conc);
}

/* (1/2 pi) int_0^{2 pi} DeltaSigma(R') cos(2 beta) dtheta, by symmetry
 * (1/pi) int_0^pi, split at theta = pi/2 and adaptive near theta = pi where
 * R' is smallest. */
double
y3o_dsigma_mis_ring(double R, double rmis, double lnM, double rho, double conc,
                    double epsrel)
{
  ring_ctx c;
  memset(&c, 0, sizeof(c));
  c.R = R, c.rmis = rmis, c.lnM = lnM, c.rho = rho, c.conc = conc;
  double e1, e2;
  double const a = y3o_gk21_adapt(ring_theta, &c, 0.0, 0.5 * Y3O_PI, epsrel,
                                  1e-300, &e1);
  double const b = y3o_gk21_adapt(ring_theta, &c, 0.5 * Y3O_PI, Y3O_PI,
                                  epsrel, 1e-300, &e2);
  return (a + b) / Y3O_PI;
}

static double
ring_outer(double r, void* v)
{
  ring_ctx* c = (ring_ctx*)v;
  double w;
  if (c->kernel == 0) {
    double const s2 = c->sig * c->sig;
    w = r / s2 * exp(-0.5 * r * r / s2);
  } else {
    w = y3o_roffset_gamma(r, c->tau);
  }
  if (w == 0.0) return 0.0;
  return w * y3o_dsigma_mis_ring(c->R, r, c->lnM, c->rho, c->conc,
                                 c->epsrel * 0.1);
}

/* Kernel-averaged miscentred Delta-Sigma: int_0^rmax P(r) ring(R, r) dr. */
static double
mis_kernel(ring_ctx* c, double a, double b)
{
  /* breakpoints at r = R (near-singular ring) */
  double e, s = 0.0;
  if (c->R > a && c->R < b) {
    s += y3o_gk21_adapt(ring_outer, c, a, c->R, c->epsrel, 1e-300, &e);
    s += y3o_gk21_adapt(ring_outer, c, c->R, b, c->epsrel, 1e-300, &e);
  } else {
    s += y3o_gk21_adapt(ring_outer, c, a, b, c->epsrel, 1e-300, &e);
  }
  return s;
}

double
y3o_dsigma_mis_gamma(double R, double tau, double lnM, double rho, double conc,
                     double epsrel)
{
  ring_ctx c;
  memset(&c, 0, sizeof(c));
  c.R = R, c.lnM = lnM, c.rho = rho, c.conc = conc, c.tau = tau;
  c.kernel = 1;
  c.epsrel = epsrel;
  return mis_kernel(&c, 0.0, 60.0 * tau);
}

double
y3o_dsigma_mis_rayleigh(double R, double sigma, double rmax, double lnM,
                        double rho, double conc, double epsrel)
{
  ring_ctx c;
  memset(&c, 0, sizeof(c));
  c.R = R, c.lnM = lnM, c.rho = rho, c.conc = conc, c.sig = sigma;
  c.kernel = 0;
  c.epsrel = epsrel;
  double const norm = 1.0 - exp(-0.5 * rmax * rmax / (sigma * sigma));
  return mis_kernel(&c, 0.0, rmax) / norm;
}

/* ---------------- golden nested integrator ------------------------------ */
/*
 * All integrals of one parameter point at once.  The integrands factor as
 *   hmf(lnM, zt) dV(zt) [Pz_bin(zt)] [Sigma_crit^-1(zt)]
 *     x int dlt MOR(lt | lbar(lnM, zt)) K_i(lt)          (K_i: lambda_obs bin)
 *     x [ (1 - f_mis) DS(R) + f_mis DS_mis(R, M; sigma_i) ]   (Delta-Sigma)
 * where the lambda_obs and z_obs integrals are done in closed form (erf and
 * EMG CDF differences; y3o_lo_lt_bin, y3o_zo_bin), and the (R_mis, theta)
 * integral by nested adaptive GK21 (y3o_dsigma_mis_rayleigh).
 *
 * The lt integral depends on (lnM, zt) only through L = ln lbar; it is
 * evaluated by adaptive GK21 at the nodes of a piecewise Chebyshev
 * interpolant in L (degree 15 on panels of width 0.1), and the
 * Delta-Sigma factor likewise in lnM (degree 15 on panels of width 0.25).
 * Outer (lnM, zt) integration: Gauss-Legendre on panels bounded by every
 * table knot (HMF, D_A, Sigma_crit^-1) and every bin edge, so each panel sees
 * a smooth integrand.
 */
#define CH_DEG 16

typedef struct {
  double a, w; /* panel start, width */
  int np;
  double* coef; /* np * nf * CH_DEG */
  int nf;
} cheb_t;

static void
cheb_nodes(double* t)
{
  for (int k = 0; k < CH_DEG; ++k)
    t[k] = cos(Y3O_PI * (k + 0.5) / CH_DEG);
}

/* coefficients from values at cheb nodes (DCT) */
static void
cheb_fit(const double* vals, double* coef)
{
  for (int j = 0; j < CH_DEG; ++j) {
    double s = 0.0;
    for (int k = 0; k < CH_DEG; ++k)
      s += vals[k] * cos(Y3O_PI * j * (k + 0.5) / CH_DEG);
    coef[j] = s * 2.0 / CH_DEG;
  }
  coef[0] *= 0.5;
}

static double
cheb_eval(const double* coef, double t)
{
  double b1 = 0.0, b2 = 0.0;
  for (int j = CH_DEG - 1; j >= 1; --j) {
    double const b0 = 2.0 * t * b1 - b2 + coef[j];
    b2 = b1;
    b1 = b0;
  }
  return t * b1 - b2 + coef[0];
}

static void
cheb_eval_all(const cheb_t* c, double x, double* out)
{
  int ip = (int)floor((x - c->a) / c->w);
  if (ip < 0) ip = 0;
  if (ip >= c->np) ip = c->np - 1;
  double const t = 2.0 * (x - (c->a + ip * c->w)) / c->w - 1.0;
  for (int f = 0; f < c->nf; ++f)
    out[f] = cheb_eval(c->coef + ((size_t)ip * c->nf + f) * CH_DEG, t);
}

typedef struct {
  const y3o_params* p;
  double L;
  const double* lo_edges;
  int ibin;
} jl_ctx;

static double
jl_integrand(double lt, void* v)
{
  jl_ctx* c = (jl_ctx*)v;
  double const s = c->p->mor_sigma_ln;
  double const d = (log(lt) - c->L) / s;
  double const mor = exp(-0.5 * d * d) * 0.3989422804014327 / (s * lt);
  return mor * y3o_lo_lt_bin(c->p, c->lo_edges[c->ibin],
                             c->lo_edges[c->ibin + 1], lt);
}

/* J_i(L) = int_1^300 MOR(lt; L) K_i(lt) dlt, split at the peak region. */
static double
jl_value(const y3o_params* p, const double* lo_edges, int ibin, double L)
{
  jl_ctx c = {p, L, lo_edges, ibin};
  double pts[8];
  int np = 0;
  pts[np++] = 1.0;
  double const cands[5] = {exp(L - 3.0 * p->mor_sigma_ln), exp(L),
                           exp(L + 3.0 * p->mor_sigma_ln), lo_edges[ibin],
                           lo_edges[ibin + 1]};
  /* sorted insertion of interior breakpoints */
  double tmp[5];
  int nt = 0;
  for (int k = 0; k < 5; ++k)
    if (cands[k] > 1.0 && cands[k] < 300.0) tmp[nt++] = cands[k];
  for (int i = 1; i < nt; ++i)
    for (int j = i; j > 0 && tmp[j] < tmp[j - 1]; --j) {
      double const t = tmp[j];
      tmp[j] = tmp[j - 1];
      tmp[j - 1] = t;
    }
  for (int k = 0; k < nt; ++k) pts[np++] = tmp[k];
  pts[np++] = 300.0;
  double s = 0.0, e;
  for (int k = 0; k + 1 < np; ++k)
    if (pts[k + 1] > pts[k])
      s += y3o_gk21_adapt(jl_integrand, &c, pts[k], pts[k + 1], 1e-12, 1e-17,
                          &e);
  return s;
}

static int golden_ngz = 4, golden_ngm = 6;

/* Gauss-Legendre orders of the golden's outer panels (convergence tests). */
void
y3o_golden_set_orders(int ngz, int ngm)
{
  golden_ngz = ngz;
  golden_ngm = ngm;
}

static int
dcmp(const void* a, const void* b)
{
  double const x = *(const double*)a, y = *(const double*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

int
y3o_golden(const y3o_params* p, const y3o_tables* t, int count_kind,
           const double* lo_edges, int n_lo, const double* zo_edges, int n_zo,
           const double* radii, int n_r, double* counts, double* dsig)
{
  y3o_params pp = *p;
  if (count_kind == Y3O_COUNT3D) pp.f_prj = 0.0;
  const y3o_table_spec* s = &t->spec;
  int const nzb = n_zo > 0 ? n_zo : 1;
  double const ln3e14 = 33.334803590584751;

  /* zt ranges per z_obs bin (y3o_make_integral) */
  double zr_lo[16], zr_hi[16];
  for (int j = 0; j < nzb; ++j) {
    zr_lo[j] = n_zo > 0 ? zo_edges[j] - 0.1 : s->z_lo;
    zr_hi[j] = n_zo > 0 ? zo_edges[j + 1] + 0.1 : s->z_hi;
  }
  double zmin = zr_lo[0], zmax = zr_hi[0];
  for (int j = 1; j < nzb; ++j) {
    if (zr_lo[j] < zmin) zmin = zr_lo[j];
    if (zr_hi[j] > zmax) zmax = zr_hi[j];
  }

  /* --- J_i(L) chebyshev over the reachable L range --- */
  double const Lmin = pp.mor_lnl0 + pp.mor_slope * (s->lnm_lo - ln3e14) +
                      pp.mor_zslope * log((1.0 + zmin) / 1.35) - 0.01;
  double const Lmax = pp.mor_lnl0 + pp.mor_slope * (s->lnm_hi - ln3e14) +
                      pp.mor_zslope * log((1.0 + zmax) / 1.35) + 0.01;
  cheb_t cj;
  cj.w = 0.1;
  cj.a = Lmin;
  cj.np = (int)ceil((Lmax - Lmin) / cj.w);
  cj.nf = n_lo;
  cj.coef = (double*)malloc(sizeof(double) * cj.np * n_lo * CH_DEG);
  double tn[CH_DEG], vals[CH_DEG];
  cheb_nodes(tn);
  for (int ip = 0; ip < cj.np; ++ip)
    for (int i = 0; i < n_lo; ++i) {
      for (int k = 0; k < CH_DEG; ++k) {
        double const L = cj.a + ip * cj.w + 0.5 * cj.w * (tn[k] + 1.0);
        vals[k] = jl_value(&pp, lo_edges, i, L);
      }
      cheb_fit(vals, cj.coef + ((size_t)ip * n_lo + i) * CH_DEG);
    }

  /* --- Delta-Sigma factor chebyshev in lnM --- */
  int const nds = n_r > 0 ? n_lo * n_r : 0;
  cheb_t cd;
  memset(&cd, 0, sizeof(cd));
  if (nds > 0) {
    double const rho = pp.omega_m * Y3O_RHOC;
    cd.w = 0.25;
    cd.a = s->lnm_lo;
    cd.np = (int)ceil((s->lnm_hi - s->lnm_lo) / cd.w - 1e-9);
    cd.nf = nds;
    cd.coef = (double*)malloc(sizeof(double) * cd.np * nds * CH_DEG);
    for (int ip = 0; ip < cd.np; ++ip)
      for (int i = 0; i < n_lo; ++i) {
        double const rl = pow(0.5 * (lo_edges[i] + lo_edges[i + 1]) / 100.0,
                              0.2);
        double const sig = pp.mis_rlambda * rl, rmax = 6.0 * sig;
        for (int r = 0; r < n_r; ++r) {
          for (int k = 0; k < CH_DEG; ++k) {
            double const lnM = cd.a + ip * cd.w + 0.5 * cd.w * (tn[k] + 1.0);
            double const cen = y3o_nfw_dsigma(radii[r], lnM, rho, 0.0);
            double const mis = y3o_dsigma_mis_rayleigh(radii[r], sig, rmax,
                                                       lnM, rho, 0.0, 1e-10);
            vals[k] = (1.0 - pp.f_mis) * cen + pp.f_mis * mis;
          }
          cheb_fit(vals,
                   cd.coef + ((size_t)ip * nds + i * n_r + r) * CH_DEG);
        }
      }
  }

  /* --- outer breakpoints --- */
  int const maxb = s->n_z + s->n_da + s->n_sci + 2 * nzb + 4;
  double* zb = (double*)malloc(sizeof(double) * maxb);
  int nb = 0;
  zb[nb++] = zmin;
  zb[nb++] = zmax;
  for (int j = 0; j < nzb; ++j) zb[nb++] = zr_lo[j], zb[nb++] = zr_hi[j];
  for (int k = 0; k < s->n_z; ++k)
    if (t->z[k] > zmin && t->z[k] < zmax) zb[nb++] = t->z[k];
  for (int k = 0; k < s->n_da; ++k)
    if (t->da_z[k] > zmin && t->da_z[k] < zmax) zb[nb++] = t->da_z[k];
  if (n_r > 0)
    for (int k = 0; k < s->n_sci; ++k)
      if (t->sci_z[k] > zmin && t->sci_z[k] < zmax) zb[nb++] = t->sci_z[k];
  qsort(zb, nb, sizeof(double), dcmp);
  {
    int m = 1;
    for (int k = 1; k < nb; ++k)
      if (zb[k] - zb[m - 1] > 1e-13) zb[m++] = zb[k];
    nb = m;
  }

  int const NGZ = golden_ngz, NGM = golden_ngm;
  double gzx[8], gzw[8], gmx[8], gmw[8];
  y3o_gl_rule(NGZ, gzx, gzw);
  y3o_gl_rule(NGM, gmx, gmw);

  int const nout = n_lo * nzb;
  double* accc = (double*)calloc(nout, sizeof(double));
  double* accd = (double*)calloc((size_t)nout * (n_r > 0 ? n_r : 1),
                                 sizeof(double));
  double* Jv = (double*)malloc(sizeof(double) * n_lo);
  double* Dv = (double*)malloc(sizeof(double) * (nds > 0 ? nds : 1));
  int const nm = s->n_lnm;
  /* precompute lnM nodes, weights, DS factors */
  int const nmn = (nm - 1) * NGM;
  double* mn = (double*)malloc(sizeof(double) * nmn);
  double* mw = (double*)malloc(sizeof(double) * nmn);
  double* mds = (double*)malloc(sizeof(double) * nmn * (nds > 0 ? nds : 1));
  for (int c = 0; c < nm - 1; ++c)
    for (int g = 0; g < NGM; ++g) {
      double const a = t->lnm[c], b = t->lnm[c + 1];
      int const k = c * NGM + g;
      mn[k] = 0.5 * (a + b) + 0.5 * (b - a) * gmx[g];
      mw[k] = 0.5 * (b - a) * gmw[g];
      if (nds > 0) cheb_eval_all(&cd, mn[k], mds + (size_t)k * nds);
    }

  for (int zp = 0; zp + 1 < nb; ++zp) {
    double const za = zb[zp], zbb = zb[zp + 1];
    double const zc = 0.5 * (za + zbb);
    int inb[16];
    for (int j = 0; j < nzb; ++j)
      inb[j] = (zc > zr_lo[j] && zc < zr_hi[j]);
    for (int gz = 0; gz < NGZ; ++gz) {
      double const zt = 0.5 * (za + zbb) + 0.5 * (zbb - za) * gzx[gz];
      double const wz = 0.5 * (zbb - za) * gzw[gz];
      double const dv = y3o_dvdz(&pp, t, zt);
      double const sci = n_r > 0 ? y3o_sci(t, zt) : 0.0;
      double pzb[16];
      for (int j = 0; j < nzb; ++j)
        pzb[j] = !inb[j] ? 0.0
                 : n_zo > 0
                   ? y3o_zo_bin(&pp, zo_edges[j], zo_edges[j + 1], zt)
                   : 1.0;
      double const lz = pp.mor_zslope * log((1.0 + zt) / 1.35);
      for (int k = 0; k < nmn; ++k) {
        double const lnM = mn[k];
        double const L = pp.mor_lnl0 + pp.mor_slope * (lnM - ln3e14) + lz;
        cheb_eval_all(&cj, L, Jv);
        double const base = wz * mw[k] * y3o_hmf(&pp, t, lnM, zt) * dv;
        for (int i = 0; i < n_lo; ++i) {
          double const bj = base * Jv[i];
          for (int j = 0; j < nzb; ++j) {
            if (!inb[j]) continue;
            double const v = bj * pzb[j];
            accc[i * nzb + j] += v;
            if (n_r > 0) {
              double const vs = v * sci;
              const double* dsk = mds + (size_t)k * nds + i * n_r;
              double* o = accd + ((size_t)(i * nzb + j)) * n_r;
              for (int r = 0; r < n_r; ++r) o[r] += vs * dsk[r];
            }
          }
        }
      }
    }
  }
  for (int q = 0; q < nout; ++q) counts[q] = accc[q];
  if (n_r > 0 && dsig)
    for (int q = 0; q < nout * n_r; ++q) dsig[q] = accd[q];

  free(cj.coef);
  free(cd.coef);
  free(zb);
  free(accc);
  free(accd);
  free(Jv);
  free(Dv);
  free(mn);
  free(mw);
  free(mds);
  return 0;
}

/* ---------------- Genz-Malik rule -------------------------------------- */
#define L2 0.35856858280031809 /* sqrt(9/70) */
#define L3 0.94868329805051380 /* sqrt(9/10) */
#define L4 0.94868329805051380 /* sqrt(9/10) */
#define L5 0.68824720161168529 /* sqrt(9/19) */

int
y3o_gm_num_nodes(int d)
{
  return 1 + 4 * d + 2 * d * (d - 1) + (1 << d);
}

/* Node j of the degree-7 rule in the enumeration used on the device:
 * [0] centre; [1, 1+2d) +-l2 e_i; [1+2d, 1+4d) +-l3 e_i;
 * [1+4d, 1+4d+2d(d-1)) +-l4 e_i +-l4 e_k (i<k); then 2^d +-l5 corners.
 * Writes the unit offsets u[d] and returns the generator class 1..5. */
static int
gm_node(int d, int j, double* u)
{
  for (int i = 0; i < d; ++i) u[i] = 0.0;
  if (j == 0) return 1;
  j -= 1;
  if (j < 2 * d) {
    u[j >> 1] = (j & 1) ? -L2 : L2;
    return 2;
  }
  j -= 2 * d;
  if (j < 2 * d) {
    u[j >> 1] = (j & 1) ? -L3 : L3;
    return 3;
  }
  j -= 2 * d;
  if (j < 2 * d * (d - 1)) {
    int pair = j >> 2, sg = j & 3, a = 0, b = 1;
    /* pair index -> (a, b), a < b, lexicographic */
    int cnt = 0;
    for (a = 0; a < d; ++a) {
      int const np = d - 1 - a;
      if (pair < cnt + np) {
        b = a + 1 + (pair - cnt);
        break;
      }
      cnt += np;
    }
    u[a] = (sg & 1) ? -L4 : L4;
    u[b] = (sg & 2) ? -L4 : L4;
    return 4;
  }
  j -= 2 * d * (d - 1);
  for (int i = 0; i < d; ++i) u[i] = ((j >> i) & 1) ? -L5 : L5;
  return 5;
}

static void
gm_weights(int d, double* w7, double* w5)
{
  double const dd = d;
  w7[1] = (12824.0 - 9120.0 * dd + 400.0 * dd * dd) / 19683.0;
  w7[2] = 980.0 / 6561.0;
  w7[3] = (1820.0 - 400.0 * dd) / 19683.0;
  w7[4] = 200.0 / 19683.0;
  w7[5] = 6859.0 / 19683.0 / (double)(1 << d);
  w5[1] = (729.0 - 950.0 * dd + 50.0 * dd * dd) / 729.0;
  w5[2] = 245.0 / 486.0;
  w5[3] = (265.0 - 100.0 * dd) / 1458.0;
  w5[4] = 25.0 / 729.0;
  w5[5] = 0.0;
}

/* Region error from the degree-7, -5 and -3 rules (device: rule_error in
 * y3g_cub_common.hpp): E1 = |R7 - R5| is the error of R5; when the rule
 * sequence converges (r = E1 / |R5 - R3| small) the error of R7 is a
 * fraction of it, E1 * clamp(c r, smin, 1); c = 0 keeps E1. */
static double
rule_error(double r7, double r5, double r3, double c, double smin)
{
  double const e1 = fabs(r7 - r5), e2 = fabs(r5 - r3);
  if (!(c > 0.0) || !(e2 > 0.0)) return e1;
  double f = c * e1 / e2;
  f = f < smin ? smin : (f > 1.0 ? 1.0 : f);
  return e1 * f;
}

double
y3o_gm_rule(double (*f)(const double*, void*), void* ctx, int d,
            const double* c, const double* hw, double* r5, double* fourth)
{
  double r3;
  return y3o_gm_rule3(f, ctx, d, c, hw, r5, &r3, fourth);
}

/* As y3o_gm_rule, also returning the degree-3 rule on the centre and the
 * +-l3 axis nodes (weights 1/(6 l3^2) and 1 - 2d/(6 l3^2)). */
double
y3o_gm_rule3(double (*f)(const double*, void*), void* ctx, int d,
             const double* c, const double* hw, double* r5, double* r3,
             double* fourth)
{
  double w7[6], w5[6], u[8], x[8];
  gm_weights(d, w7, w5);
  double const w33 = 10.0 / 54.0, w31 = 1.0 - 2.0 * d * w33;
  int const n = y3o_gm_num_nodes(d);
  double s7 = 0.0, s5 = 0.0, s3 = 0.0, f0 = 0.0, f2p[8], f3p[8];
  for (int j = 0; j < n; ++j) {
    int const g = gm_node(d, j, u);
    for (int i = 0; i < d; ++i) x[i] = c[i] + hw[i] * u[i];
    double const fv = f(x, ctx);
    s7 += w7[g] * fv;
    s5 += w5[g] * fv;
    if (g == 1) s3 += w31 * fv;
    if (g == 3) s3 += w33 * fv;
    if (j == 0) f0 = fv;
    if (g == 2) {
      int const a = (j - 1) >> 1;
      if ((j - 1) & 1)
        f2p[a] += fv;
      else
        f2p[a] = fv;
    }
    if (g == 3) {
      int const a = (j - 1 - 2 * d) >> 1;
      if ((j - 1 - 2 * d) & 1)
        f3p[a] += fv;
      else
        f3p[a] = fv;
    }
  }
  double vol = 1.0;
  for (int i = 0; i < d; ++i) vol *= 2.0 * hw[i];
  if (fourth)
    for (int i = 0; i < d; ++i)
      fourth[i] = fabs(f2p[i] - 2.0 * f0 - (f3p[i] - 2.0 * f0) / 7.0);
  *r5 = s5 * vol;
  *r3 = s3 * vol;
  return s7 * vol;
}

/* ---------------- adaptive cubature ------------------------------------ */
/*
 * Region algorithm shared with the device kernel (csrc/y3g_cubature.hip):
 *   1. initial partition: k pieces per axis (4 for d <= 3, else 2);
 *   2. per region: R7, R5, err = |R7 - R5|, split axis = argmax fourth
 *      difference (ties -> lowest axis; all ~0 -> widest relative edge);
 *   3. loop: I = sum R7, E = sum err; stop when E <= max(epsabs, epsrel|I|);
 *      otherwise split every region with err > E / N (threshold doubled
 *      until at most max_split_per_step are selected), each along its axis
 *      at the midpoint;
 *      child A replaces the parent, child B is appended;
 *   4. consistency: if |R7_A + R7_B - R7_parent| > err_parent, both
 *      children's errors are scaled by that ratio;
 *   5. stop with status 1 when max_eval or max_regions would be exceeded.
 */
typedef struct {
  double (*f)(const double*, void*);
  void* ctx;
  int d;
  double lows[8], highs[8];
  double err_c, err_smin;
  double sb; /* smoothing bound, relative to |I| (y3o_smooth_bound) */
  int am;    /* lnM axis of a synthetic kind (init_split_lnm), -1 none */
} cub_prob;

typedef struct {
  double lo[8], hi[8];
  double est, err, raw; /* raw = the region's own |R7 - R5| */
  int axis;
} region_t;

/* Two-level scale of a split region's children (device: two_level_scale in
 * y3g_cub_common.hpp): clamp(c d / e, smin, 1), d = |R7(A) + R7(B) -
 * R7(parent)|, e = the parent's |R7 - R5|; c = 0 leaves them unscaled. */
static double
two_level_scale(double d, double e, double c, double smin)
{
  if (!(c > 0.0) || !(e > 0.0)) return 1.0;
  double const s = c * d / e;
  return s < smin ? smin : (s > 1.0 ? 1.0 : s);
}

/* Budget extension (device: ext_budget in y3g_cub_common.hpp): a call
 * limit lim that would cut the next refinement step while the reported
 * error err is above ratio x tol is doubled, up to ext_eval; the integral
 * then refines as a run with its final limit from the start. */
static int
ext_budget(long long lim, long long ext_eval, double ratio, double err, double tol)
{
  return ratio > 0.0 && ext_eval > lim && err > ratio * tol;
}

static void
eval_region(const cub_prob* P, region_t* r)
{
  double c[8], hw[8], fourth[8], r5;
  for (int i = 0; i < P->d; ++i) {
    c[i] = 0.5 * (r->lo[i] + r->hi[i]);
    hw[i] = 0.5 * (r->hi[i] - r->lo[i]);
  }
  double r3;
  r->est = y3o_gm_rule3(P->f, P->ctx, P->d, c, hw, &r5, &r3, fourth);
  r->err = rule_error(r->est, r5, r3, P->err_c, P->err_smin);
  r->raw = r->err;
  int ax = 0;
  double best = fourth[0];
  for (int i = 1; i < P->d; ++i)
    if (fourth[i] > best) best = fourth[i], ax = i;
  if (!(best > 0.0)) {
    double bw = -1.0;
    for (int i = 0; i < P->d; ++i) {
      double const rel = (r->hi[i] - r->lo[i]) / (P->highs[i] - P->lows[i]);
      if (rel > bw) bw = rel, ax = i;
    }
  }
  r->axis = ax;
}

/* Error bin: binary exponent of E minus that of err (0 = largest), 63 for
 * err == 0 (device: err_bin in y3g_kernels.hip). */
static int
err_bin(double err, double E, int eE)
{
  if (!(err > 0.0)) return 63;
  int e;
  (void)frexp(err, &e);
  int const b = eE - e;
  return b < 0 ? 0 : (b > 62 ? 62 : b);
}

/* err / E in 2^-40 fixed point, qs = 2^40 / E computed once per step
 * (device: err_q / err_qscale) */
static unsigned long long
err_q(double err, double qs)
{
  double const q = floor(err * qs);
  return q > 0.0 ? (unsigned long long)q : 0ull;
}

/* Split bin B: first bin where the cumulative error mass reaches one half
 * and first where the cumulative count reaches max(64, N/8) (capped by N and
 * room); bins < B are split (up to room), then bin B in index order. */
static void
select_bins(const int* hc, const unsigned long long* hq, int N, int room,
            int* B, int* lim0, int* lim1)
{
  unsigned long long const qa = (unsigned long long)(0.5 * 1099511627776.0);
  int km = N / 8 > 64 ? N / 8 : 64;
  if (km > room) km = room;
  if (km > N) km = N;
  int Ba = 63, Bk = 63, cc = 0;
  unsigned long long cq = 0;
  for (int b = 0; b < 64; ++b) {
    cq += hq[b];
    cc += hc[b];
    if (Ba == 63 && cq >= qa) Ba = b;
    if (Bk == 63 && cc >= km) Bk = b;
  }
  *B = Ba > Bk ? Ba : Bk;
  int below = 0;
  for (int b = 0; b < *B; ++b) below += hc[b];
  *lim0 = below < room ? below : room;
  int const inb = *B < 63 ? hc[*B] : 0;
  *lim1 = room - *lim0 < inb ? room - *lim0 : inb;
}

static void
cub_run(const cub_prob* Pin, const y3o_cub_opts* o, y3o_result* res)
{
  cub_prob Pc = *Pin;
  Pc.err_c = o->err_c;
  Pc.err_smin = o->err_smin;
  const cub_prob* P = &Pc;
  int const d = P->d;
  int const nn = y3o_gm_num_nodes(d);
  /* the safety factor applies with the smoothed HMF only (P->sb > 0) */
  double const escale = (P->sb > 0.0 && o->err_scale > 0.0) ? o->err_scale : 1.0;
  int cap = o->max_regions > 0 ? o->max_regions : 8192;
  int const kmax = o->max_split_per_step > 0 ? o->max_split_per_step : 64;
  region_t* R = (region_t*)malloc(sizeof(region_t) * cap);
  int* selidx = (int*)malloc(sizeof(int) * (kmax > 0 ? kmax : 1));
  long long neval = 0;
  int n = 1;
  /* initial partition: k pieces per axis (k = init_split, or 4 for d <= 3
   * and 2 otherwise); region r has base-k digits q_a (axis 0 fastest) and
   * spans [L + q w, L + (q+1) w] with w = (H - L) / k, the end pieces ending
   * exactly at L and H */
  /* pieces per axis: k everywhere, init_split_lnm along a synthetic
   * kind's lnM axis (device: init_pieces in y3g_cub_common.hpp) */
  int kp[8];
  {
    int const ks = o->init_split > 0 ? o->init_split : (d <= 3 ? 4 : 2);
    long long nn0 = 1;
    for (int a = 0; a < d; ++a) {
      kp[a] = (o->init_split_lnm > 0 && a == P->am) ? o->init_split_lnm : ks;
      nn0 *= kp[a];
    }
    if (nn0 > cap) {
      for (int a = 0; a < d; ++a) kp[a] = 1;
      nn0 = 1;
    }
    n = (int)nn0;
  }
  for (int r = 0; r < n; ++r) {
    int rr = r;
    for (int a = 0; a < d; ++a) {
      int const ks = kp[a];
      int const q = rr % ks;
      rr /= ks;
      double const L = P->lows[a], H = P->highs[a];
      double const w = (H - L) / ks;
      R[r].lo[a] = q == 0 ? L : L + q * w;
      R[r].hi[a] = q == ks - 1 ? H : L + (q + 1) * w;
    }
  }
  for (int k = 0; k < n; ++k) eval_region(P, &R[k]);
  neval += (long long)n * nn;
  int status = 1;
  double I = 0.0, E = 0.0;
  long long lim = o->max_eval;
  long long const ext_eval = o->ext_regions > 0 ? 2LL * o->ext_regions * nn : 0;
  for (;;) {
    I = 0.0;
    E = 0.0;
    for (int k = 0; k < n; ++k) I += R[k].est, E += R[k].err;
    double const tol = fmax(o->eps_abs, o->eps_rel * fabs(I));
    if (escale * E + P->sb * fabs(I) <= tol) {
      status = 0;
      break;
    }
    /* at most kmax splits, within the capacity and the call limit */
    int room = (kmax < cap - n) ? kmax : cap - n;
    long long left = (lim - neval) / (2LL * nn);
    while (left < room && ext_budget(lim, ext_eval, o->ext_ratio, escale * E + P->sb * fabs(I), tol)) {
      lim = 2 * lim < ext_eval ? 2 * lim : ext_eval;
      left = (lim - neval) / (2LL * nn);
    }
    if (left < room) room = (int)left;
    if (room <= 0) break;
    int eE, hc[64], B, lim0, lim1;
    unsigned long long hq[64];
    (void)frexp(E, &eE);
    memset(hc, 0, sizeof(hc));
    memset(hq, 0, sizeof(hq));
    double const qs = 1099511627776.0 / E;
    for (int k = 0; k < n; ++k) {
      int const b = err_bin(R[k].err, E, eE);
      hc[b] += 1;
      hq[b] += err_q(R[k].err, qs);
    }
    select_bins(hc, hq, n, room, &B, &lim0, &lim1);
    int const ns = lim0 + lim1;
    if (ns <= 0 || neval + 2LL * ns * nn > lim) break;
    /* ordered selection: bins < B (up to lim0), then bin B (up to lim1) */
    int m = 0;
    for (int k = 0; k < n && m < lim0; ++k)
      if (err_bin(R[k].err, E, eE) < B) selidx[m++] = k;
    for (int k = 0; k < n && m < lim0 + lim1; ++k)
      if (err_bin(R[k].err, E, eE) == B) selidx[m++] = k;
    int const n0 = n;
    for (int j = 0; j < ns; ++j) {
      int const k = selidx[j];
      region_t par = R[k], a = par, b = par;
      int const ax = par.axis;
      double const s = 0.5 * (par.lo[ax] + par.hi[ax]);
      a.hi[ax] = s;
      b.lo[ax] = s;
      eval_region(P, &a);
      eval_region(P, &b);
      double const delta = fabs(a.est + b.est - par.est);
      double const sc = two_level_scale(delta, par.raw, o->tl_c, o->tl_smin);
      a.err *= sc;
      b.err *= sc;
      double const se = a.err + b.err;
      if (delta > se) {
        if (se > 0.0) {
          double const f = delta / se;
          a.err *= f;
          b.err *= f;
        } else {
          a.err = b.err = 0.5 * delta;
        }
      }
      R[k] = a;
      R[n0 + j] = b;
    }
    n = n0 + ns;
    neval += 2LL * ns * nn;
  }
  res->value = I;
  res->error = escale * E + P->sb * fabs(I);
  res->prob = 0.0;
  res->neval = neval;
  res->nregions = n;
  res->status = status;
  free(R);
  free(selidx);
}

void
y3o_cubature_fn(double (*f)(const double*, void*), void* ctx, int ndim,
                const double* lows, const double* highs,
                const y3o_cub_opts* o, y3o_result* r)
{
  cub_prob P;
  memset(&P, 0, sizeof(P));
  P.f = f;
  P.ctx = ctx;
  P.d = ndim;
  P.am = -1;
  for (int i = 0; i < ndim; ++i) P.lows[i] = lows[i], P.highs[i] = highs[i];
  cub_run(&P, o, r);
}

typedef struct {
  const y3o_params* p;
  const y3o_tables* t;
  const y3o_integral* ig;
} ig_ctx;

static double
ig_eval(const double* x, void* v)
{
  ig_ctx* c = (ig_ctx*)v;
  return y3o_integrand_mapped(c->p, c->t, c->ig, x);
}

void
y3o_cubature(const y3o_params* p, const y3o_tables* t, const y3o_integral* ig,
             const y3o_cub_opts* o, y3o_result* r)
{
  y3o_params pp = *p;
  if (ig->kind == Y3O_COUNT3D) pp.f_prj = 0.0;
  y3o_tables tt = *t;
  tt.smooth_hmf = o->hmf_smooth;
  ig_ctx c = {&pp, &tt, ig};
  cub_prob P;
  memset(&P, 0, sizeof(P));
  P.f = ig_eval;
  P.ctx = &c;
  P.d = ig->ndim;
  P.sb = y3o_smooth_bound(ig->kind, o->hmf_smooth && t->spec.n_lnm >= 6 && t->spec.n_z >= 6);
  P.am = ig->kind == Y3O_COUNT3D ? 2 : ig->kind == Y3O_COUNT5D ? 4 : 3;
  for (int i = 0; i < ig->ndim; ++i)
    P.lows[i] = ig->lows[i], P.highs[i] = ig->highs[i];
  cub_run(&P, o, r);
}

/* ---------------- Philox4x32-10 --------------------------------------- */
void
y3o_philox(uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t const p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t const p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t const hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t const hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0, out[1] = c1, out[2] = c2, out[3] = c3;
}

/* ---------------- VEGAS with m-Cubes stratification ------------------- */
/*
 * Shared with the device kernels (csrc/y3g_vegas.hip):
 *   g = floor((calls/2)^(1/d)) cubes per axis (>= 1), p = max(2, calls/g^d)
 *   samples per cube; sample (it, cube, j) draws its d uniforms from
 *   Philox4x32-10 with key = (seed_lo, seed_hi) and counter
 *   (cube, j, it, block), 53-bit doubles from word pairs.
 *   Importance grid: nbins per axis, Lepage refinement with smoothing and
 *   damping exponent alpha after every iteration.  Iterations < itskip only
 *   adapt; later ones are combined with inverse-variance weights; the
 *   reported error is Y3O_VEGAS_NSIG standard deviations, inflated by
 *   sqrt(chi2/dof) when that exceeds 1, so that it bounds the true error
 *   with ~99.7% probability (a 1-sigma estimate does not bound it).
 */
#define Y3O_VEGAS_NSIG 3.0
static void
vegas_refine(double* xg, const double* hist, int nb, double alpha)
{
  double d[1024], r[1024], xn[1025];
  double sum = 0.0;
  if (nb == 1) {
    d[0] = hist[0];
  } else {
    d[0] = 0.5 * (hist[0] + hist[1]);
    d[nb - 1] = 0.5 * (hist[nb - 2] + hist[nb - 1]);
    for (int b = 1; b < nb - 1; ++b)
      d[b] = (hist[b - 1] + hist[b] + hist[b + 1]) / 3.0;
  }
  for (int b = 0; b < nb; ++b) sum += d[b];
  if (!(sum > 0.0)) return;
  double rs = 0.0;
  for (int b = 0; b < nb; ++b) {
    double const q = d[b] / sum;
    r[b] = (q > 0.0 && q < 1.0) ? pow((q - 1.0) / log(q), alpha) : 0.0;
    rs += r[b];
  }
  if (!(rs > 0.0)) return;
  double const target = rs / nb;
  int k = -1;
  double dr = 0.0;
  xn[0] = 0.0;
  for (int i = 1; i < nb; ++i) {
    while (dr < target && k < nb - 1) {
      ++k;
      dr += r[k];
    }
    dr -= target;
    xn[i] = xg[k + 1] - (xg[k + 1] - xg[k]) * dr / r[k];
  }
  xn[nb] = 1.0;
  for (int i = 0; i <= nb; ++i) xg[i] = xn[i];
}

static void vegas_run(double (*f)(const double*, void*), void* ctx, int d,
                      const double* lows, const double* highs,
                      const y3o_vegas_opts* o, double sb, y3o_result* res);

void
y3o_vegas_fn(double (*f)(const double*, void*), void* ctx, int d,
             const double* lows, const double* highs,
             const y3o_vegas_opts* o, y3o_result* res)
{
  vegas_run(f, ctx, d, lows, highs, o, 0.0, res);
}

/* VEGAS with sb x |I| added to the error it tests and reports (the
 * smoothing bound of the synthetic kinds, y3o_smooth_bound). */
static void
vegas_run(double (*f)(const double*, void*), void* ctx, int d,
          const double* lows, const double* highs,
          const y3o_vegas_opts* o, double sb, y3o_result* res)
{
  int const nb = o->nbins;
  double* xg = (double*)malloc(sizeof(double) * d * (nb + 1));
  double* hist = (double*)malloc(sizeof(double) * d * nb);
  for (int k = 0; k < d; ++k)
    for (int b = 0; b <= nb; ++b) xg[k * (nb + 1) + b] = (double)b / nb;
  /* calls per iteration, fewer when the call limit could not hold all
   * iterations (as the device kernel) */
  long long calls = o->calls_per_iter;
  long long const per_it = o->max_eval / (o->itmax > 0 ? o->itmax : 1);
  if (per_it < calls) calls = per_it;
  if (calls < 2) calls = 2;
  int g = (int)floor(pow(calls / 2.0, 1.0 / d));
  if (g < 1) g = 1;
  long long ncubes = 1;
  for (int k = 0; k < d; ++k) ncubes *= g;
  long long p = calls / ncubes;
  if (p < 2) p = 2;
  double range[8], vol = 1.0;
  for (int k = 0; k < d; ++k) range[k] = highs[k] - lows[k], vol *= range[k];
  uint32_t const key[2] = {(uint32_t)o->seed, (uint32_t)(o->seed >> 32)};
  double wsum = 0.0, isum = 0.0, chisum = 0.0, i2sum = 0.0;
  int nacc = 0, status = 1;
  long long neval = 0;
  double I = 0.0, sig = 0.0;
  for (int it = 0; it < o->itmax; ++it) {
    memset(hist, 0, sizeof(double) * d * nb);
    double sI = 0.0, sV = 0.0;
    for (long long c = 0; c < ncubes; ++c) {
      int ck[8];
      long long cc = c;
      for (int k = 0; k < d; ++k) ck[k] = (int)(cc % g), cc /= g;
      double s1 = 0.0, s2 = 0.0;
      for (long long j = 0; j < p; ++j) {
        uint32_t w[16];
        int const nblk = (2 * d + 3) / 4;
        for (int blk = 0; blk < nblk; ++blk) {
          uint32_t ctr[4] = {(uint32_t)c, (uint32_t)j, (uint32_t)it,
                             (uint32_t)blk};
          y3o_philox(ctr, key, w + 4 * blk);
        }
        double x[8], jac = vol;
        int bins[8];
        for (int k = 0; k < d; ++k) {
          double const u = ((double)(w[2 * k] >> 5) * 67108864.0 +
                            (double)(w[2 * k + 1] >> 6)) *
                           (1.0 / 9007199254740992.0);
          double const y = (ck[k] + u) / g;
          double const z = y * nb;
          int b = (int)z;
          if (b > nb - 1) b = nb - 1;
          double const* xk = xg + k * (nb + 1);
          double const wb = xk[b + 1] - xk[b];
          x[k] = lows[k] + (xk[b] + (z - b) * wb) * range[k];
          jac *= nb * wb;
          bins[k] = b;
        }
        double const fx = f(x, ctx) * jac;
        s1 += fx;
        s2 += fx * fx;
        for (int k = 0; k < d; ++k) hist[k * nb + bins[k]] += fx * fx;
      }
      double const m = s1 / p;
      double const v = (s2 - s1 * m) / ((double)p * (p - 1));
      sI += m;
      sV += v > 0.0 ? v : 0.0;
    }
    neval += ncubes * p;
    double const Ii = sI / ncubes;
    double const Vi = sV / ((double)ncubes * ncubes);
    if (it >= o->itskip) {
      double const wi = Vi > 0.0 ? 1.0 / Vi : 1e300;
      wsum += wi;
      isum += wi * Ii;
      i2sum += wi * Ii * Ii;
      ++nacc;
      I = isum / wsum;
      chisum = i2sum - wsum * I * I; /* sum w (I_i - I)^2 */
      sig = sqrt(1.0 / wsum);
      double const dof = nacc - 1;
      double const chi = dof > 0 ? chisum / dof : 0.0;
      double const sigr = Y3O_VEGAS_NSIG * (chi > 1.0 ? sig * sqrt(chi) : sig) + sb * fabs(I);
      if (sigr <= fmax(o->eps_abs, o->eps_rel * fabs(I))) {
        status = 0;
        break;
      }
    }
    if (neval + ncubes * p > o->max_eval) break;
    if (it + 1 < o->itmax)
      for (int k = 0; k < d; ++k)
        vegas_refine(xg + k * (nb + 1), hist + k * nb, nb, o->alpha);
  }
  double const dof = nacc - 1;
  double const chi = dof > 0 ? chisum / dof : 0.0;
  res->value = I;
  res->error = Y3O_VEGAS_NSIG * (chi > 1.0 ? sig * sqrt(chi) : sig) + sb * fabs(I);
  res->prob = chi;
  res->neval = neval;
  res->nregions = 0;
  res->status = status;
  free(xg);
  free(hist);
}

void
y3o_vegas(const y3o_params* p, const y3o_tables* t, const y3o_integral* ig,
          const y3o_vegas_opts* o, y3o_result* r)
{
  y3o_params pp = *p;
  if (ig->kind == Y3O_COUNT3D) pp.f_prj = 0.0;
  y3o_tables tt = *t;
  tt.smooth_hmf = o->hmf_smooth;
  ig_ctx c = {&pp, &tt, ig};
  vegas_run(ig_eval, &c, ig->ndim, ig->lows, ig->highs, o,
            y3o_smooth_bound(ig->kind, o->hmf_smooth && t->spec.n_lnm >= 6 && t->spec.n_z >= 6), r);
}

/* ---------------- reference integrator test functions ------------------ */
/* integration.test.cc:33-91 and ExampleScalarIntegrand.test.cc:9-60 */
double
y3o_test_fn(int which, const double* x, int ndim)
{
  (void)ndim;
  switch (which) {
    case 0: { /* sf1 = 0.5 on [0,1] */
      double const s = sin(Y3O_PI * x[0]);
      return s * s;
    }
    case 1: /* sf2 = 3xy(x+y), = 1 on unit square, -1276 on [-4,7]x[-11,-3] */
      return 3.0 * x[0] * x[1] * (x[0] + x[1]);
    case 2: /* sf6 = 4.14199 on the unit 6-cube */
      return 44.0 * (sin(Y3O_PI * x[0] * x[1]) / x[0]) * x[3] * x[4] *
             pow(x[2], x[4]) * sin(Y3O_PI * x[5] * x[5]);
    case 3: /* constant */
      return 1.0;
    case 4: /* ExampleScalarIntegrand: x / radius + (y - sigma_8)^2 with
               radius 2.5, sigma_8 0.75: 0.345833 on the unit square,
               346.688 on [-3,2]x[4,7] */
      return x[0] / 2.5 + (x[1] - 0.75) * (x[1] - 0.75);
    default:
      return 0.0;
  }
}
