// evaluators.hpp -- the classes of the CosmoSIS evaluator modules (one
// module library per class, NumCountsSel.cc).  An evaluator does a fixed
// quadrature of its own per sample and is queried per grid point; the
// concept is the reference's CosmoSISScalarEvaluatorModule integrand
// (src/utils/CosmoSISScalarEvaluatorModule.hh:9-26): grid_t / grid_point_t,
// n_outputs, a constructor from the config block, set_sample, evaluate,
// module_label, output_sections and make_grid_points.  Here the constructor
// also takes the process's device runtime, and set_sample runs the
// quadrature on the device through liby3g.
#ifndef Y3G_EVALUATORS_HPP
#define Y3G_EVALUATORS_HPP

#include "y3g_module.hpp"

// NumCountsSel -- CosmoSIS module: cluster number counts per (lambda_obs,
// z_obs) bin through a precomputed selection function S_b(lnM, zt), on fixed
// Gauss-Legendre nodes (the reference's NumCountsSelGL,
// src/models/n_operator_sel_gl_t.hh:171-216 over SelGLCore, :74-168):
//
//   N_b = sum_k w_k sum_q w_q dV/dz(z_q) Omega(z_q) hmf(lnM_k, z_q) S_b(lnM_k, z_q)
//
// with n_lnm x n_z nodes over [lnm_low, lnm_high] x [zt_low, zt_high] and
// Omega(z) the effective survey area.  All bins of a sample are one device
// launch (y3g_selgl_counts); evaluate() picks bin bin_index.
//
// Config (<label> = NumCountsSel):
//   n_lnm (default 96), n_z (default 64), zt_low, zt_high, lnm_low, lnm_high
//   bin_index                                 -- grid (wall of numbers)
//   table_lnm_low / _high, table_z_low / _high -- optional, the window of the
//                                                device HMF table built from
//                                                the synthetic physics
// Sample:
//   cosmological_parameters/omega_m (+ the optional y3g_synthetic and
//   cluster_abundance keys of SyntheticSample), mass_function/{m_h, z,
//   dndlnmh} and distances/{z, d_a} when present (else built on the device),
//   sel_function/{lnM [n_lnm], z [n_z], S_stack [n_bins, n_z, n_lnm]}
//   (sel_function_t.hh:5-7).
// Output: numcountssel/vals, one value per grid point.

class NumCountsSel : public y3g_cosmosis::SyntheticSample {
public:
  using grid_t = y3g_cosmosis::grid_t<1>;
  using grid_point_t = grid_t::value_type;
  static constexpr std::size_t n_outputs = 1;

  NumCountsSel(cosmosis::DataBlock& cfg, y3g_runtime* rt) : rt_(rt)
  {
    std::string const L = module_label();
    gl_.n_lnm = cfg.has_val(L, "n_lnm") ? cfg.view<int>(L, "n_lnm") : 96;
    gl_.n_z = cfg.has_val(L, "n_z") ? cfg.view<int>(L, "n_z") : 64;
    gl_.zt_lo = cfg.view<double>(L, "zt_low");
    gl_.zt_hi = cfg.view<double>(L, "zt_high");
    gl_.lnm_lo = cfg.view<double>(L, "lnm_low");
    gl_.lnm_hi = cfg.view<double>(L, "lnm_high");
    if (gl_.n_lnm < 1 || gl_.n_z < 1 || gl_.n_lnm > Y3G_SELGL_MAX_NODES ||
        gl_.n_z > Y3G_SELGL_MAX_NODES)
      throw std::runtime_error(L + ": n_lnm and n_z must lie in [1, " +
                               std::to_string(Y3G_SELGL_MAX_NODES) + "]");
    y3g_table_spec s;
    y3g_spec_wide(&s);
    s.lnm_lo = y3g_cosmosis::optional_double(cfg, L, "table_lnm_low", s.lnm_lo);
    s.lnm_hi = y3g_cosmosis::optional_double(cfg, L, "table_lnm_high", s.lnm_hi);
    s.z_lo = y3g_cosmosis::optional_double(cfg, L, "table_z_low", s.z_lo);
    s.z_hi = y3g_cosmosis::optional_double(cfg, L, "table_z_high", s.z_hi);
    set_spec(s);
  }

  void
  set_sample(cosmosis::DataBlock& s)
  {
    SyntheticSample::set_sample(s);
    tables_.ensure(rt_, *this);
    using doubles = std::vector<double>;
    auto const lnm = s.view<doubles>("sel_function", "lnM");
    auto const z = s.view<doubles>("sel_function", "z");
    auto const nd = s.view<cosmosis::ndarray<double>>("sel_function", "S_stack");
    auto const ext = nd.extents();
    if (ext.size() != 3)
      throw std::runtime_error("sel_function/S_stack must be 3-D (n_bins, n_z, n_lnm)");
    if (ext[1] != z.size() || ext[2] != lnm.size())
      throw std::runtime_error("sel_function: S_stack extents mismatch");
    int const nb = int(ext[0]);
    std::vector<double> const planes(nd.begin(), nd.end());
    if (y3g_selfn_upload(rt_, 0, lnm.data(), int(lnm.size()), z.data(), int(z.size()),
                         planes.data(), nb) != Y3G_OK)
      throw std::runtime_error(std::string("sel_function upload: ") + y3g_last_error());
    norm_.assign(std::size_t(nb), 0.0);
    if (y3g_selgl_counts(rt_, &gl_, 1, nb, norm_.data()) != Y3G_OK)
      throw std::runtime_error(std::string("selection counts: ") + y3g_last_error());
  }

  std::array<double, n_outputs>
  evaluate(grid_point_t const& pt) const
  {
    int const b = static_cast<int>(pt[0]);
    if (b < 0 || static_cast<std::size_t>(b) >= norm_.size())
      throw std::out_of_range("NumCountsSelGL: bin_index outside sel_function/S_stack range");
    return {norm_[std::size_t(b)]};
  }

  static char const* module_label() { return "NumCountsSel"; }
  static std::array<char const*, n_outputs> output_sections() { return {"numcountssel"}; }
  static grid_t
  make_grid_points(cosmosis::DataBlock& cfg)
  {
    return y3g_cosmosis::make_grid_points_wall_of_numbers<1>(cfg, module_label(),
                                                             {"bin_index"});
  }

private:
  y3g_runtime* rt_;
  y3g_selgl_spec gl_{};
  y3g_cosmosis::TableCache tables_;
  std::vector<double> norm_;
};

#endif
