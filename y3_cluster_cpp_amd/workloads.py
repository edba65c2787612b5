"""Synthetic workloads of the integration path.

- the prior-box stream: parameter points drawn from Omega_m in [0.2, 0.4],
  HMF amplitude in [0.7, 0.9], MOR slope in [0.7, 1.1], sigma_ln in
  [0.1, 0.4], f_mis in [0, 0.4] (Philox key 0x59334331), each carrying 192
  integrals -- 12 five-dimensional number counts (4 lambda_obs x 3 z_obs
  bins) and 180 six-dimensional miscentred Delta-Sigma integrals (the 12
  bins x 15 radii log-spaced in [0.1, 30] Mpc/h) -- with per-integral
  eps_rel log-uniform in [1e-5, 1e-3] (Philox key 0x59334331 + 1);
- the serving stream: requests arriving as a Poisson process from a mock
  random-walk MCMC over the same box, 70% asking for the 12 counts only and
  30% for counts + Delta-Sigma (192 integrals).  The chain moves between the
  stream's first points (those with FP64 oracle values), proposing one of
  the nearest neighbours of its current point and accepting with a
  Metropolis rule on a smooth mock likelihood, so every served integral can
  be checked.
"""
import numpy as np

LO_EDGES = [20.0, 30.0, 45.0, 60.0, 300.0]
ZO_EDGES = [0.20, 0.35, 0.50, 0.65]
RADII = list(np.logspace(np.log10(0.1), np.log10(30.0), 15))
SEED = 0x59334331
BOX_LO = np.array([0.2, 0.7, 0.7, 0.1, 0.0])
BOX_HI = np.array([0.4, 0.9, 1.1, 0.4, 0.4])
N_COUNTS = 12
N_PER_POINT = 192


def prior_points(n, seed=SEED):
    rng = np.random.Generator(np.random.Philox(key=seed))
    u = rng.random((n, 5))
    return BOX_LO + u * (BOX_HI - BOX_LO)


def eps_stream(n_points, seed=SEED):
    """Per-integral eps_rel, log-uniform in [1e-5, 1e-3], 192 per point."""
    rng = np.random.Generator(np.random.Philox(key=seed + 1))
    return 10.0 ** rng.uniform(-5.0, -3.0, (n_points, N_PER_POINT))


def point_params(y3, pt):
    om, amp, slope, sln, fmis = pt
    return y3.default_params(omega_m=om, hmf_amp=amp, mor_slope=slope, mor_sigma_ln=sln,
                             f_mis=fmis)


def point_integrals(y3, params, spec, k, eps_row, algorithm, counts_only=False):
    """The 192 (or 12) integrals of one parameter point, in the golden
    values' order: counts (lambda_obs bin major), then Delta-Sigma (bin
    major, radius minor)."""
    igs = []
    n = 0
    for i in range(4):
        for j in range(3):
            igs.append(y3.make_integral(params, spec, y3.COUNT5D, LO_EDGES[i], LO_EDGES[i + 1],
                                        ZO_EDGES[j], ZO_EDGES[j + 1], point=k,
                                        eps_rel=float(eps_row[n]), algorithm=algorithm))
            n += 1
    if counts_only:
        return igs
    for i in range(4):
        for j in range(3):
            for r in RADII:
                igs.append(y3.make_integral(params, spec, y3.DSIGMA6D, LO_EDGES[i],
                                            LO_EDGES[i + 1], ZO_EDGES[j], ZO_EDGES[j + 1], r,
                                            point=k, eps_rel=float(eps_row[n]),
                                            algorithm=algorithm))
                n += 1
    return igs


def count_ok(res, ref, eps, cap=1e-4):
    """Counted: |I - I_oracle| <= eps |I_oracle|, <= cap |I_oracle|, and the
    reported error bounds the true error."""
    err = np.abs(res["value"] - ref)
    rel = err / np.abs(ref)
    ok = (rel <= eps) & (rel <= cap) & (res["error"] >= err)
    return ok, rel


def mcmc_requests(points, n_requests, rate, seed=SEED + 2, frac_full=0.3, k_nn=6):
    """Arrival times (s) and (point index, full?) of a mock random-walk MCMC.

    Requests arrive as a Poisson process of the given rate.  The chain's
    proposals walk the k_nn-nearest-neighbour graph of `points` (box
    coordinates scaled to [0, 1]); each proposal is one request, 30% of them
    for the full 192 integrals and the rest for the 12 counts only.
    """
    rng = np.random.Generator(np.random.Philox(key=seed))
    u = (np.asarray(points) - BOX_LO) / (BOX_HI - BOX_LO)
    d2 = ((u[:, None, :] - u[None, :, :]) ** 2).sum(-1)
    nn = np.argsort(d2, axis=1)[:, 1:k_nn + 1]
    centre = np.full(5, 0.5)

    def loglike(i):
        return -0.5 * float(((u[i] - centre) ** 2).sum()) / 0.15 ** 2

    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    cur = int(rng.integers(len(points)))
    reqs = []
    for _ in range(n_requests):
        prop = int(nn[cur, rng.integers(k_nn)])
        reqs.append((prop, bool(rng.random() < frac_full)))
        if np.log(rng.random()) < loglike(prop) - loglike(cur):
            cur = prop
    return t, reqs


# Fixed-GL number counts with a selection function: the reference's default
# grids (sel_function.py: n_lnm 40, n_z_shared 20, 12 bins) and GL nodes
# (n_operator_sel_gl_t.hh: n_lnm 96, n_z 64).
SEL_BINS, SEL_NLNM, SEL_NZ = 12, 40, 20
SEL_GL = (30.5, 35.0, 0.15, 0.75, 96, 64)  # lnm_lo, lnm_hi, zt_lo, zt_hi, n_lnm, n_z


def selection_stack(rng, n_bins=SEL_BINS, nm=SEL_NLNM, nz=SEL_NZ):
    """A smooth synthetic sel_function/{lnM, z, S_stack}: lnM linear in
    [30, 35.5], z linear in [0.1, 0.8], S_stack [n_bins, n_z, n_lnm] with
    richness bins (b % 4) as sigmoids in lnM and redshift bins (b // 4) as
    Gaussian windows in z, the shape of S_ij = S_i(lnM, z) K_j(z)."""
    lnm = np.linspace(30.0, 35.5, nm)
    z = np.linspace(0.1, 0.8, nz)
    M, Z = np.meshgrid(lnm, z)
    out = np.empty((n_bins, nz, nm))
    for b in range(n_bins):
        m0 = 31.5 + 0.5 * (b % 4) + rng.uniform(-0.1, 0.1)
        zc = 0.25 + 0.15 * (b // 4)
        out[b] = (1 / (1 + np.exp(-(M - m0) / 0.3))) * np.exp(-0.5 * ((Z - zc) / 0.08) ** 2)
    return lnm, z, out


# The real-data Delta-Sigma integrand (SigmaMiscentY1MortCUDAScalarIntegrand):
# synthetic DES inputs of the reference's shapes.
DES_HEAD = (0.15,                                   # roffset_tau (Mpc/h)
            1.2, 0.1, 3.0, 0.9,                     # LO_LC alpha, a, b, R_lambda
            10 ** 11.2, 10 ** 12.4, 0.75, 0.2, 0.1, 0.3)  # MOR_DES A, B, alpha, sigma, eps, z_p


def des_synthetic(n_r=30, n_m=13, n_z=10):
    """{roffset_tau, LO_LC, MOR_DES, SIG_MAX} for y3g_des_upload: SIG_MAX
    tables on n_r log-spaced radii in [0.05, 40] Mpc/h, n_m lnM in [30, 36],
    n_z z in [0.1, 1.0]: a one-halo term A(M) r^-0.9 exp(-r / 4 r_200) and a
    two-halo term 0.08 r^-0.7 with bias (1 + (M / 1e14)^0.5)(1 + z)^0.5, the
    reference's SIG_MAX inputs in shape (deltasigma/sigma_1 [n_lnM][n_r],
    sigma_2 [n_z][n_r], bias [n_lnM][n_z]).  The reference's real tables
    are 1000 radii x 300 masses (data/deltasigma_r_perp.txt,
    deltasigma_m_h.txt)."""
    r = np.logspace(np.log10(0.05), np.log10(40.0), n_r)
    m = np.linspace(30.0, 36.0, n_m)
    z = np.linspace(0.1, 1.0, n_z)
    M = np.exp(m)[:, None]
    r200 = 1.6 * (M / 1e15) ** (1 / 3)
    s1 = 60.0 * (M / 1e14) ** 0.6 * r[None, :] ** -0.9 * np.exp(-r[None, :] / (4 * r200))
    s2 = np.broadcast_to(0.08 * r[None, :] ** -0.7, (len(z), len(r)))
    bias = (1 + (np.exp(m)[:, None] / 1e14) ** 0.5) * (1 + z[None, :]) ** 0.5
    return np.concatenate([list(DES_HEAD), [len(r), len(m), len(z)], r, m, z,
                           s1.reshape(-1), np.ascontiguousarray(s2).reshape(-1),
                           bias.reshape(-1)])


def des_volumes():
    """The reference demo's binning for the real-data Delta-Sigma module:
    4 lambda_obs x 3 z_obs bins x 15 radii log-spaced in [0.1, 30] Mpc/h
    (180 integrals), each over lc in [1, 300], zt within 0.1 of its z_obs
    bin, lnM in [29.93, 35.69], rmis in [0, 1.5] Mpc/h, theta in [0, 2 pi].
    Returns [(lows, highs, (zo_low, zo_high, radius))]."""
    out = []
    for a in range(4):
        for b in range(3):
            zlo, zhi = ZO_EDGES[b], ZO_EDGES[b + 1]
            lows = [LO_EDGES[a], 1.0, zlo - 0.1, 29.9336, 0.0, 0.0]
            highs = [LO_EDGES[a + 1], 300.0, zhi + 0.1, 35.6901, 1.5, 2 * np.pi]
            for R in RADII:
                out.append((lows, highs, (zlo, zhi, R)))
    return out
