// integrands.hpp -- the integrand classes of the CosmoSIS modules (one
// module library per class, NumberCounts3DIntegrand.cc etc.).  Each class
// models the reference's module-integrand concept: grid_t / grid_point_t,
// a constructor from the config block, set_sample, set_grid_point,
// module_label, make_integration_volumes and make_grid_points (reference:
// src/utils/CosmoSISSICUDAModule.hh:99-110 and the modules under
// src/modules/*/*_cuda_scalarintegrand.cu).  In place of the device
// operator() they describe each integral for liby3g, whose device integrand
// evaluates it.
#ifndef Y3G_INTEGRANDS_HPP
#define Y3G_INTEGRANDS_HPP

#include "y3g_module.hpp"

// NumberCounts3DIntegrand -- CosmoSIS module: cluster number counts in one
// lambda_obs bin, integrated over (lc, zt, lnm) with the lambda_obs bin done
// in closed form (lc is the true richness).
//
// Per integral:  int dlnM dzt dlc  n(lnM, zt) P(lc | lnM, zt) dV/dz(zt)
//                    x [P(lo_high > lo > lo_low | lc)]
// The integrand is the device integrand Y3G_COUNT3D of liby3g.  The module
// concept mirrors the reference's count modules, e.g.
// src/modules/mass_mort_y1/NC_cent_y1_mort_cuda_scalarintegrand.cu:164-175
// (integrand body) and :190-196 (volumes).
//
// Config (<label> = NumberCounts3DIntegrand):
//   algorithm, eps_rel, eps_abs, max_eval, use_cartesian_product
//   lc_low/lc_high, zt_low/zt_high, lnm_low/lnm_high  -- volumes
//   lo_low, lo_high                                   -- grid (lambda_obs bin)

class NumberCounts3DIntegrand : public y3g_cosmosis::SyntheticSample {
public:
  using grid_t = y3g_cosmosis::grid_t<2>;
  using grid_point_t = grid_t::value_type;

  explicit NumberCounts3DIntegrand(cosmosis::DataBlock&) {}
  void set_grid_point(grid_point_t pt) { bin_ = pt; }
  y3g_integral
  describe(y3g_cosmosis::volume_t const& v) const
  {
    return describe_kind(Y3G_COUNT3D, v, bin_.data(), 2);
  }
  static char const* module_label() { return "NumberCounts3DIntegrand"; }
  static std::vector<y3g_cosmosis::volume_t>
  make_integration_volumes(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_integration_volumes_wall_of_numbers(
      cfg, module_label(), {"lc", "zt", "lnm"});
  }
  static grid_t
  make_grid_points(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_grid_points_wall_of_numbers<2>(
      cfg, module_label(), {"lo_low", "lo_high"});
  }

private:
  grid_point_t bin_{};
};

// NumberCounts5DIntegrand -- CosmoSIS module: cluster number counts in a
// (lambda_obs, z_obs) bin, integrated over (lo, zo, lc, zt, lnm).
//
// Per integral:  int dlo dzo dlc dzt dlnM  n(lnM, zt) P(lc | lnM, zt)
//                    P(lo | lc) P(zo | zt) dV/dz(zt)
// with P(lo | lc) the Gaussian + projection-tail (EMG) richness scatter and
// P(zo | zt) the photometric-redshift Gaussian.  The integrand is the device
// integrand Y3G_COUNT5D of liby3g; the module concept mirrors the reference's
// src/modules/mass_mort_y1/NC_miscent_y1_mort_cuda_scalarintegrand.cu:153-193.
// The bins are the integration volumes themselves, so there is no grid.
//
// Config (<label> = NumberCounts5DIntegrand):
//   algorithm, eps_rel, eps_abs, max_eval, use_cartesian_product
//   lo_*, zo_*, lc_*, zt_*, lnm_* (_low / _high)      -- volumes

class NumberCounts5DIntegrand : public y3g_cosmosis::SyntheticSample {
public:
  using grid_t = y3g_cosmosis::grid_t<0>;
  using grid_point_t = grid_t::value_type;

  explicit NumberCounts5DIntegrand(cosmosis::DataBlock&) {}
  void set_grid_point(grid_point_t) {}
  y3g_integral
  describe(y3g_cosmosis::volume_t const& v) const
  {
    return describe_kind(Y3G_COUNT5D, v, nullptr, 0);
  }
  static char const* module_label() { return "NumberCounts5DIntegrand"; }
  static std::vector<y3g_cosmosis::volume_t>
  make_integration_volumes(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_integration_volumes_wall_of_numbers(
      cfg, module_label(), {"lo", "zo", "lc", "zt", "lnm"});
  }
  static grid_t
  make_grid_points(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_grid_points_wall_of_numbers<0>(cfg, module_label(), {});
  }
};

// SigmaMiscent6DIntegrand -- CosmoSIS module: the stacked, miscentred
// excess surface density Delta-Sigma(R) of a (lambda_obs, z_obs) bin,
// integrated over (lo, lc, zt, lnm, rmis, theta) with the z_obs bin in
// closed form.
//
// Per integral:  int dlo dlc dzt dlnM drmis dtheta/2pi  n P(lc|lnM,zt)
//                  P(lo|lc) [P(zo bin|zt)] dV/dz Sigma_crit^-1(zt) P(rmis)
//                  [(1 - f_mis) DS_NFW(R) + f_mis DS_t(R; rmis, theta)]
// where DS_t is the NFW Delta-Sigma at R' = sqrt(R^2 + rmis^2 + 2 R rmis
// cos theta) projected on the tangential direction.  The integrand is the
// device integrand Y3G_DSIGMA6D of liby3g.  Variable order, volumes and grid
// follow the reference module
// src/modules/sigma_mort_y1/sigma_miscent_y1_mort_cuda_scalarintegrand.cu:
// operator() (lo, lc, zt, lnM, rmis, theta) :90-95 and :160-182, volumes
// :190-203, grid (zo_low, zo_high, radii) :205-215.
//
// Config (<label> = SigmaMiscent6DIntegrand):
//   algorithm, eps_rel, eps_abs, max_eval, use_cartesian_product
//   lo_*, lc_*, zt_*, lnm_*, rmis_*, theta_* (_low / _high) -- volumes
//   zo_low, zo_high, radii                                  -- grid

class SigmaMiscent6DIntegrand : public y3g_cosmosis::SyntheticSample {
public:
  using grid_t = y3g_cosmosis::grid_t<3>;
  using grid_point_t = grid_t::value_type;

  explicit SigmaMiscent6DIntegrand(cosmosis::DataBlock&) {}
  void set_grid_point(grid_point_t pt) { gp_ = pt; }
  y3g_integral
  describe(y3g_cosmosis::volume_t const& v) const
  {
    return describe_kind(Y3G_DSIGMA6D, v, gp_.data(), 3);
  }
  static char const* module_label() { return "SigmaMiscent6DIntegrand"; }
  static std::vector<y3g_cosmosis::volume_t>
  make_integration_volumes(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_integration_volumes_wall_of_numbers(
      cfg, module_label(), {"lo", "lc", "zt", "lnm", "rmis", "theta"});
  }
  static grid_t
  make_grid_points(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_grid_points_wall_of_numbers<3>(
      cfg, module_label(), {"zo_low", "zo_high", "radii"});
  }

private:
  grid_point_t gp_{};
};

// ----------------------------------------------------- real-data (DES) modules
//
// The reference's Y1 real-data modules under their own labels, keys and
// variables; their integrands are the device kinds Y3G_*_DES (FP64 batched
// cubature or m-Cubes).  Sample inputs, read as the reference's functors
// read them in set_sample:
//   ROFFSET_t     cluster_abundance/roffset_tau               (miscentred)
//   LO_LC_t       cluster_abundance/LO_LC_alpha, LO_LC_a, LO_LC_b,
//                 LO_LC_R_lambda                              (miscentred)
//   MOR_DES_t     cluster_abundance/mor_A, mor_B, mor_alpha, mor_sigma,
//                 mor_epsilon, z_mor_pivot
//   SIG_MAX       deltasigma/{r_sigma_deltasigma, lnM, sigma_1, sigma_2,
//                 bias}, matter_power_lin/z                   (Delta-Sigma)
//   HMF_t, DV_DO_DZ_t  the sample's mass_function / distances tables when
//                 present (else the synthetic tables); OMEGA_Z_DES and
//                 INT_ZO_ZT_DES_t have no inputs.
// Config (<label> = the module's): algorithm, eps_rel, eps_abs, max_eval,
// use_cartesian_product, the volumes and the grid below; precision fp64.

class DesSample : public y3g_cosmosis::SyntheticSample {
public:
  static constexpr bool fp64_only = true;

  void
  upload(y3g_runtime* rt) const
  {
    if (y3g_des_upload(rt, 0, des_.data(), int(des_.size())) != Y3G_OK)
      throw std::runtime_error(std::string("DES data upload: ") + y3g_last_error());
  }

protected:
  // The {roffset, LO_LC, MOR_DES[, SIG_MAX]} block of y3g_des_upload; the
  // centred modules, which read no offset keys, carry zeros there.
  void
  read_des(cosmosis::DataBlock& sample, bool miscent, bool sigma)
  {
    using y3g_cosmosis::view_as_double;
    using doubles = std::vector<double>;
    std::string const A = "cluster_abundance";
    des_.clear();
    for (char const* k : {"roffset_tau", "LO_LC_alpha", "LO_LC_a", "LO_LC_b", "LO_LC_R_lambda"})
      des_.push_back(miscent ? view_as_double(sample, A, k) : 0.0);
    for (char const* k : {"mor_A", "mor_B", "mor_alpha", "mor_sigma", "mor_epsilon",
                          "z_mor_pivot"})
      des_.push_back(view_as_double(sample, A, k));
    if (!sigma) return;
    auto const r = sample.view<doubles>("deltasigma", "r_sigma_deltasigma");
    auto const m = sample.view<doubles>("deltasigma", "lnM");
    auto const z = sample.view<doubles>("matter_power_lin", "z");
    auto const s1 = sample.view<cosmosis::ndarray<double>>("deltasigma", "sigma_1");
    auto const s2 = sample.view<cosmosis::ndarray<double>>("deltasigma", "sigma_2");
    auto const bi = sample.view<cosmosis::ndarray<double>>("deltasigma", "bias");
    if (s1.size() != r.size() * m.size() || s2.size() != r.size() * z.size() ||
        bi.size() != z.size() * m.size())
      throw std::runtime_error("deltasigma/sigma_1, sigma_2, bias must hold n_r n_lnM, "
                               "n_r n_z and n_z n_lnM values");
    des_.push_back(double(r.size()));
    des_.push_back(double(m.size()));
    des_.push_back(double(z.size()));
    des_.insert(des_.end(), r.begin(), r.end());
    des_.insert(des_.end(), m.begin(), m.end());
    des_.insert(des_.end(), z.begin(), z.end());
    des_.insert(des_.end(), s1.begin(), s1.end());
    des_.insert(des_.end(), s2.begin(), s2.end());
    des_.insert(des_.end(), bi.begin(), bi.end());
  }

  std::vector<double> des_;
};

// SigmaMiscentY1MortCUDAScalarIntegrand -- miscentred Delta-Sigma over
// (lo, lc, zt, lnm, rmis, theta), grid (zo_low, zo_high, radii)
// (sigma_miscent_y1_mort_cuda_scalarintegrand.cu:161-182, 190-215).
class SigmaMiscentY1MortCUDAScalarIntegrand : public DesSample {
public:
  using grid_t = y3g_cosmosis::grid_t<3>;
  using grid_point_t = grid_t::value_type;

  explicit SigmaMiscentY1MortCUDAScalarIntegrand(cosmosis::DataBlock&) {}
  void set_grid_point(grid_point_t pt) { gp_ = pt; }
  void
  set_sample(cosmosis::DataBlock& sample)
  {
    SyntheticSample::set_sample(sample);
    read_des(sample, true, true);
  }
  y3g_integral
  describe(y3g_cosmosis::volume_t const& v) const
  {
    return describe_kind(Y3G_DSIGMA6D_DES, v, gp_.data(), 3);
  }
  static char const* module_label() { return "SigmaMiscentY1MortCUDAScalarIntegrand"; }
  static std::vector<y3g_cosmosis::volume_t>
  make_integration_volumes(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_integration_volumes_wall_of_numbers(
      cfg, module_label(), {"lo", "lc", "zt", "lnm", "rmis", "theta"});
  }
  static grid_t
  make_grid_points(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_grid_points_wall_of_numbers<3>(
      cfg, module_label(), {"zo_low", "zo_high", "radii"});
  }

private:
  grid_point_t gp_{};
};

// SigmaCentY1MortCUDAScalarIntegrand -- centred Delta-Sigma over
// (lo, zt, lnm), grid (zo_low, zo_high, radii)
// (sigma_cent_y1_mort_cuda_scalarintegrand.cu:155-166, 177-195).
class SigmaCentY1MortCUDAScalarIntegrand : public DesSample {
public:
  using grid_t = y3g_cosmosis::grid_t<3>;
  using grid_point_t = grid_t::value_type;

  explicit SigmaCentY1MortCUDAScalarIntegrand(cosmosis::DataBlock&) {}
  void set_grid_point(grid_point_t pt) { gp_ = pt; }
  void
  set_sample(cosmosis::DataBlock& sample)
  {
    SyntheticSample::set_sample(sample);
    read_des(sample, false, true);
  }
  y3g_integral
  describe(y3g_cosmosis::volume_t const& v) const
  {
    return describe_kind(Y3G_DSIGMA3D_DES, v, gp_.data(), 3);
  }
  static char const* module_label() { return "SigmaCentY1MortCUDAScalarIntegrand"; }
  static std::vector<y3g_cosmosis::volume_t>
  make_integration_volumes(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_integration_volumes_wall_of_numbers(cfg, module_label(),
                                                                  {"lo", "zt", "lnm"});
  }
  static grid_t
  make_grid_points(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_grid_points_wall_of_numbers<3>(
      cfg, module_label(), {"zo_low", "zo_high", "radii"});
  }

private:
  grid_point_t gp_{};
};

// NCCentY1MortCUDAScalarIntegrand -- centred counts over (lo, zt, lnm),
// grid (zo_low, zo_high) (NC_cent_y1_mort_cuda_scalarintegrand.cu:165-175,
// 192-227).
class NCCentY1MortCUDAScalarIntegrand : public DesSample {
public:
  using grid_t = y3g_cosmosis::grid_t<2>;
  using grid_point_t = grid_t::value_type;

  explicit NCCentY1MortCUDAScalarIntegrand(cosmosis::DataBlock&) {}
  void set_grid_point(grid_point_t pt) { gp_ = pt; }
  void
  set_sample(cosmosis::DataBlock& sample)
  {
    SyntheticSample::set_sample(sample);
    read_des(sample, false, false);
  }
  y3g_integral
  describe(y3g_cosmosis::volume_t const& v) const
  {
    return describe_kind(Y3G_COUNT3D_DES, v, gp_.data(), 2);
  }
  static char const* module_label() { return "NCCentY1MortCUDAScalarIntegrand"; }
  static std::vector<y3g_cosmosis::volume_t>
  make_integration_volumes(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_integration_volumes_wall_of_numbers(cfg, module_label(),
                                                                  {"lo", "zt", "lnm"});
  }
  static grid_t
  make_grid_points(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_grid_points_wall_of_numbers<2>(cfg, module_label(),
                                                             {"zo_low", "zo_high"});
  }

private:
  grid_point_t gp_{};
};

// NCMiscentY1MortCUDAScalarIntegrand -- miscentred counts over
// (lo, lc, zt, lnm, rmis), grid (zo_low, zo_high)
// (NC_miscent_y1_mort_cuda_scalarintegrand.cu:153-164, 172-194).
class NCMiscentY1MortCUDAScalarIntegrand : public DesSample {
public:
  using grid_t = y3g_cosmosis::grid_t<2>;
  using grid_point_t = grid_t::value_type;

  explicit NCMiscentY1MortCUDAScalarIntegrand(cosmosis::DataBlock&) {}
  void set_grid_point(grid_point_t pt) { gp_ = pt; }
  void
  set_sample(cosmosis::DataBlock& sample)
  {
    SyntheticSample::set_sample(sample);
    read_des(sample, true, false);
  }
  y3g_integral
  describe(y3g_cosmosis::volume_t const& v) const
  {
    return describe_kind(Y3G_COUNT5D_DES, v, gp_.data(), 2);
  }
  static char const* module_label() { return "NCMiscentY1MortCUDAScalarIntegrand"; }
  static std::vector<y3g_cosmosis::volume_t>
  make_integration_volumes(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_integration_volumes_wall_of_numbers(
      cfg, module_label(), {"lo", "lc", "zt", "lnm", "rmis"});
  }
  static grid_t
  make_grid_points(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_grid_points_wall_of_numbers<2>(cfg, module_label(),
                                                             {"zo_low", "zo_high"});
  }

private:
  grid_point_t gp_{};
};

// numberCountsCard -- counts over (lnLo, lnLt, zt, lnm), grid (zo_low,
// zo_high) (gt_card_gpu/numberCounts.cu:126-235): the observed-richness
// transfer INT_LC_LT_DES_t2 from the reference's table files, read from the
// working directory as the reference reads them (int_lc_lt_des_t_lt_bins.txt,
// int_lc_lt_des_t_zt_bins.txt, int_lc_lt_des_t_lbd_bin0..3.txt; values in
// file order), and MOR_DES_LOG_t from cluster_abundance/mor_logMmin,
// mor_logRatio, mor_alpha, mor_sigma, mor_epsilon, z_mor_pivot.  Config also
// takes do_cartesian_product_of_bins (bool, default false), as the
// reference's constructor does.
class numberCountsCard : public DesSample {
public:
  using grid_t = y3g_cosmosis::grid_t<2>;
  using grid_point_t = grid_t::value_type;

  explicit numberCountsCard(cosmosis::DataBlock& cfg)
  {
    // an absent key takes the default (the reference insists on success; a
    // DataBlock may report the default as DBS_USED_DEFAULT)
    auto const rc =
      cfg.get_val(module_label(), "do_cartesian_product_of_bins", false, bins_cartesian_);
    if (rc != DBS_SUCCESS && rc != DBS_USED_DEFAULT)
      throw std::runtime_error("numberCountsCard failed to find do_cartesian_product_of_bins\n");
  }
  void set_grid_point(grid_point_t pt) { gp_ = pt; }
  void
  set_sample(cosmosis::DataBlock& sample)
  {
    SyntheticSample::set_sample(sample);
    using y3g_cosmosis::view_as_double;
    std::string const A = "cluster_abundance";
    des_.assign(5, 0.0);
    for (char const* k : {"mor_logMmin", "mor_logRatio", "mor_alpha", "mor_sigma", "mor_epsilon",
                          "z_mor_pivot"})
      des_.push_back(view_as_double(sample, A, k));
    if (lclt_.empty()) lclt_ = read_lclt_tables_();
    des_.insert(des_.end(), lclt_.begin(), lclt_.end());
  }
  y3g_integral
  describe(y3g_cosmosis::volume_t const& v) const
  {
    return describe_kind(Y3G_COUNT4D_CARD, v, gp_.data(), 2);
  }
  static char const* module_label() { return "numberCountsCard"; }
  static std::vector<y3g_cosmosis::volume_t>
  make_integration_volumes(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_integration_volumes_wall_of_numbers(
      cfg, module_label(), {"lnLo", "lnLt", "zt", "lnm"});
  }
  static grid_t
  make_grid_points(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_grid_points_wall_of_numbers<2>(cfg, module_label(),
                                                             {"zo_low", "zo_high"});
  }

private:
  static std::vector<double>
  read_numbers_(std::string const& name)
  {
    std::ifstream in(name);
    if (!in) throw std::runtime_error("numberCountsCard: cannot open " + name);
    std::vector<double> v;
    double x;
    while (in >> x) v.push_back(x);
    return v;
  }
  // {n_lt, n_zt, lt knots, zt knots, four tables} as y3g_des_upload takes them
  static std::vector<double>
  read_lclt_tables_()
  {
    auto const lt = read_numbers_("int_lc_lt_des_t_lt_bins.txt");
    auto const zt = read_numbers_("int_lc_lt_des_t_zt_bins.txt");
    std::vector<double> out{double(lt.size()), double(zt.size())};
    out.insert(out.end(), lt.begin(), lt.end());
    out.insert(out.end(), zt.begin(), zt.end());
    for (int b = 0; b < 4; ++b) {
      auto const t = read_numbers_("int_lc_lt_des_t_lbd_bin" + std::to_string(b) + ".txt");
      if (t.size() != lt.size() * zt.size())
        throw std::runtime_error("numberCountsCard: int_lc_lt_des_t_lbd_bin" +
                                 std::to_string(b) + ".txt holds " + std::to_string(t.size()) +
                                 " values, not n_lt n_zt");
      out.insert(out.end(), t.begin(), t.end());
    }
    return out;
  }

  grid_point_t gp_{};
  bool bins_cartesian_ = false;
  std::vector<double> lclt_;
};

// ---------------------------------------------- Costanzi-2026 P[X] modules
//
// bSelMargP1GPU, bSelMargI1GPU, bSelMargI2GPU, bSelMargJGPU: the P[X]
// operator integrands P_operator_gpu<W> (src/models/p_operator_gpu_t.cuh:
// 118-299, weights :327-366) over (lt, zt, lnm), grid (zo_low, zo_high,
// lambda_bin); their device integrands are Y3G_POP_* (FP64).  Sample inputs,
// read as P_operator_gpu::set_sample reads them (:150-163):
//   HMF_t, DV_DO_DZ_t      the sample's mass_function / distances tables
//   MOR_SHIFTED_POISSON_t  cluster_mor/{log10_Mmin, log10_M1, alpha,
//                          sigma_lambda, epsilon, z_pivot}, else
//                          cluster_abundance/{mor_logMmin, mor_logRatio,
//                          mor_alpha, mor_sigma, mor_epsilon, z_mor_pivot}
//                          (mor_shifted_poisson_t.cuh:76-101)
//   haloModel/{lnM, z, bias}, xi_nl/{r, z, xi_nl}, distances/{z, d_c},
//   cosmological_parameters/h0
// Config: algorithm, eps_rel, eps_abs, max_eval, use_cartesian_product;
// the volume as scalars lt_low, lt_high, zt_low, zt_high, lnm_low, lnm_high
// (:262-276); the grid zo_low, zo_high, lambda_bin (:280-299).
struct PopP1 {
  static constexpr int kind = Y3G_POP_P1;
  static char const* label() { return "bSelMargP1GPU"; }
};
struct PopI1 {
  static constexpr int kind = Y3G_POP_I1;
  static char const* label() { return "bSelMargI1GPU"; }
};
struct PopI2 {
  static constexpr int kind = Y3G_POP_I2;
  static char const* label() { return "bSelMargI2GPU"; }
};
struct PopJ {
  static constexpr int kind = Y3G_POP_J;
  static char const* label() { return "bSelMargJGPU"; }
};

template <class W>
class PopIntegrand : public DesSample {
public:
  using grid_t = y3g_cosmosis::grid_t<3>;
  using grid_point_t = grid_t::value_type;

  explicit PopIntegrand(cosmosis::DataBlock&) {}
  void set_grid_point(grid_point_t pt) { gp_ = pt; }
  void
  set_sample(cosmosis::DataBlock& sample)
  {
    using y3g_cosmosis::view_as_double;
    using doubles = std::vector<double>;
    SyntheticSample::set_sample(sample);
    des_.assign(11, 0.0);
    des_[0] = view_as_double(sample, "cosmological_parameters", "h0");
    if (sample.has_val("cluster_mor", "log10_Mmin")) {
      int k = 5;
      for (char const* n : {"log10_Mmin", "log10_M1", "alpha", "sigma_lambda", "epsilon",
                            "z_pivot"})
        des_[k++] = view_as_double(sample, "cluster_mor", n);
    } else {
      std::string const A = "cluster_abundance";
      double const lmin = view_as_double(sample, A, "mor_logMmin");
      des_[5] = lmin;
      des_[6] = lmin + view_as_double(sample, A, "mor_logRatio");
      des_[7] = view_as_double(sample, A, "mor_alpha");
      des_[8] = view_as_double(sample, A, "mor_sigma");
      des_[9] = view_as_double(sample, A, "mor_epsilon");
      des_[10] = view_as_double(sample, A, "z_mor_pivot");
    }
    auto const cz = sample.view<doubles>("distances", "z");
    auto const cd = sample.view<doubles>("distances", "d_c");
    auto const bm = sample.view<doubles>("haloModel", "lnM");
    auto const bz = sample.view<doubles>("haloModel", "z");
    auto const bt = sample.view<cosmosis::ndarray<double>>("haloModel", "bias");
    auto const xr = sample.view<doubles>("xi_nl", "r");
    auto const xz = sample.view<doubles>("xi_nl", "z");
    auto const xt = sample.view<cosmosis::ndarray<double>>("xi_nl", "xi_nl");
    if (cd.size() != cz.size() || bt.size() != bm.size() * bz.size() ||
        xt.size() != xr.size() * xz.size())
      throw std::runtime_error(std::string(W::label()) +
                               ": distances/d_c, haloModel/bias and xi_nl/xi_nl must hold "
                               "n_z, n_z n_lnM and n_z n_r values");
    for (double n : {double(cz.size()), double(bm.size()), double(bz.size()),
                     double(xr.size()), double(xz.size())})
      des_.push_back(n);
    des_.insert(des_.end(), cz.begin(), cz.end());
    des_.insert(des_.end(), cd.begin(), cd.end());
    des_.insert(des_.end(), bm.begin(), bm.end());
    des_.insert(des_.end(), bz.begin(), bz.end());
    des_.insert(des_.end(), bt.begin(), bt.end());
    des_.insert(des_.end(), xr.begin(), xr.end());
    des_.insert(des_.end(), xz.begin(), xz.end());
    des_.insert(des_.end(), xt.begin(), xt.end());
  }
  y3g_integral
  describe(y3g_cosmosis::volume_t const& v) const
  {
    return describe_kind(W::kind, v, gp_.data(), 3);
  }
  static char const* module_label() { return W::label(); }
  static std::vector<y3g_cosmosis::volume_t>
  make_integration_volumes(cosmosis::DataBlock& cfg)
  {
    char const* L = module_label();
    y3g_cosmosis::volume_t v;
    v.lows = {cfg.view<double>(L, "lt_low"), cfg.view<double>(L, "zt_low"),
              cfg.view<double>(L, "lnm_low")};
    v.highs = {cfg.view<double>(L, "lt_high"), cfg.view<double>(L, "zt_high"),
               cfg.view<double>(L, "lnm_high")};
    return {v};
  }
  static grid_t
  make_grid_points(cosmosis::DataBlock& cfg)
  {
    char const* L = module_label();
    auto const a = y3g_cosmosis::get_vector_double(cfg, L, "zo_low");
    auto const b = y3g_cosmosis::get_vector_double(cfg, L, "zo_high");
    auto const c = y3g_cosmosis::get_vector_double(cfg, L, "lambda_bin");
    if (b.size() != a.size() || c.size() != a.size())
      throw std::runtime_error(std::string(L) + ": grid arrays must have same length");
    grid_t g;
    g.names = {"zo_low", "zo_high", "lambda_bin"};
    for (std::size_t i = 0; i < a.size(); ++i) g.points.push_back({a[i], b[i], c[i]});
    return g;
  }

private:
  grid_point_t gp_{};
};

using bSelMargP1GPU = PopIntegrand<PopP1>;
using bSelMargI1GPU = PopIntegrand<PopI1>;
using bSelMargI2GPU = PopIntegrand<PopI2>;
using bSelMargJGPU = PopIntegrand<PopJ>;

#endif
