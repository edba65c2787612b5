// Host-side tests of the CosmoSIS module layer (no GPU needed): the
// DataBlock subset, the config readers, volumes and grid points, the
// integral list and the result layout of the modules, and setup() failing
// cleanly.  Modelled on the reference's module test
// src/modules/ExampleScalar/test_ExampleScalarIntegrationModule.cc:14-47
// and the datablock_reader / make_grid_points behaviour it relies on.
#include "integrands.hpp"

#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <functional>
#include <string>

namespace {
  int failures = 0;
  int checks = 0;

#define CHECK(cond)                                                            \
  do {                                                                         \
    ++checks;                                                                  \
    if (!(cond)) {                                                             \
      ++failures;                                                              \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
    }                                                                          \
  } while (0)

  template <class E>
  bool
  throws(std::function<void()> f, std::string const& needle = "")
  {
    try {
      f();
    }
    catch (E const& e) {
      return needle.empty() || std::string(e.what()).find(needle) != std::string::npos;
    }
    catch (...) {
      return false;
    }
    return false;
  }

  bool
  close(double a, double b, double tol = 1e-14)
  {
    return std::fabs(a - b) <= tol * std::max(1.0, std::fabs(b));
  }

  using namespace y3g_cosmosis;
  using cosmosis::DataBlock;
  using cosmosis::ndarray;

  void
  test_datablock()
  {
    DataBlock b;
    CHECK(b.put_val("Sec", "X", 1.5) == DBS_SUCCESS);
    CHECK(b.put_val("sec", "x", 2.5) == DBS_NAME_ALREADY_EXISTS);
    double x = 0;
    CHECK(b.get_val("SEC", "x", x) == DBS_SUCCESS && x == 1.5);
    int i = 0;
    CHECK(b.get_val("sec", "x", i) == DBS_WRONG_VALUE_TYPE);
    CHECK(b.get_val("nosec", "x", x) == DBS_SECTION_NOT_FOUND);
    CHECK(b.get_val("sec", "nox", x) == DBS_NAME_NOT_FOUND);
    CHECK(b.replace_val("sec", "x", 3.5) == DBS_SUCCESS && b.view<double>("sec", "x") == 3.5);
    CHECK(b.replace_val("sec", "y", 3.5) == DBS_NAME_NOT_FOUND);
    CHECK(throws<DataBlock::BadDataBlockAccess>([&] { b.view<double>("sec", "y"); }));
    double d = 0;
    CHECK(b.get_val("sec", "y", 7.0, d) == DBS_USED_DEFAULT && d == 7.0);
    b.put_val("sec", "v", std::vector<double>{1, 2, 3});
    auto nd = b.view<ndarray<double>>("sec", "v");
    CHECK(nd.ndims() == 1 && nd.size() == 3 && nd(2) == 3.0);
    ndarray<int> m(std::vector<int>{1, 2, 3, 4, 5, 6}, {2, 3});
    CHECK(m(1, 0) == 4 && m(0, 2) == 3);
    b.put_val("sec", "m", m);
    std::vector<int> notvec;
    CHECK(b.get_val("sec", "m", notvec) == DBS_WRONG_VALUE_TYPE);  // 2-d
    b.put_val("sec", "s", std::string("pagani"));
    CHECK(b.view<std::string>("sec", "s") == "pagani");
    b.put_val("sec", "flag", true);
    CHECK(b.view<bool>("sec", "flag"));
    CHECK(b.has_section("SEC") && !b.has_section("other"));
  }

  void
  test_readers()
  {
    DataBlock b;
    b.put_val("s", "vd", std::vector<double>{0.5, 1.5});
    b.put_val("s", "d", 2.5);
    b.put_val("s", "vi", std::vector<int>{3, 4});
    b.put_val("s", "i", 7);
    b.put_val("s", "str", std::string("x"));
    CHECK((get_vector_double(b, "s", "vd") == std::vector<double>{0.5, 1.5}));
    CHECK((get_vector_double(b, "s", "d") == std::vector<double>{2.5}));
    CHECK((get_vector_double(b, "s", "vi") == std::vector<double>{3, 4}));
    CHECK((get_vector_double(b, "s", "i") == std::vector<double>{7}));
    CHECK(throws<std::runtime_error>([&] { get_vector_double(b, "s", "str"); },
                                     "could be coerced to type vector<double>"));
    CHECK(view_as_double(b, "s", "i") == 7.0 && view_as_double(b, "s", "d") == 2.5);
    CHECK(throws<std::runtime_error>([&] { view_as_double(b, "s", "missing"); }));
    CHECK(optional_double(b, "s", "missing", 4.0) == 4.0);
  }

  void
  test_volumes_and_grids()
  {
    DataBlock c;
    std::string const L = "lab";
    c.put_val(L, "a_low", std::vector<double>{0, 1});
    c.put_val(L, "a_high", std::vector<double>{1, 2});
    c.put_val(L, "b_low", std::vector<double>{10, 20});
    c.put_val(L, "b_high", std::vector<double>{11, 21});
    auto v = make_integration_volumes_wall_of_numbers(c, L, {"a", "b"});
    CHECK(v.size() == 2);
    CHECK(v[1].lows == (std::vector<double>{1, 20}) && v[1].highs == (std::vector<double>{2, 21}));
    c.put_val(L, "c_low", std::vector<double>{0});
    c.put_val(L, "c_high", std::vector<double>{1});
    CHECK(throws<std::runtime_error>(
      [&] { make_integration_volumes_wall_of_numbers(c, L, {"a", "c"}); }, "bad, bad user!"));
    // cartesian product: first variable slowest
    auto cp = make_integration_volumes_cartesian_product(c, L, {"a", "b"});
    CHECK(cp.size() == 4);
    CHECK(cp[1].lows == (std::vector<double>{0, 20}) && cp[2].lows == (std::vector<double>{1, 10}));
    CHECK(cp[3].highs == (std::vector<double>{2, 21}));
    c.put_val(L, "g1", std::vector<double>{1, 2, 3});
    c.put_val(L, "g2", std::vector<double>{4, 5, 6});
    c.put_val(L, "g3", std::vector<double>{7, 8});
    auto g = make_grid_points_wall_of_numbers<2>(c, L, {"g1", "g2"});
    CHECK(g.size() == 3 && g.points[2][0] == 3 && g.points[2][1] == 6);
    CHECK(throws<std::invalid_argument>(
      [&] { make_grid_points_wall_of_numbers<2>(c, L, {"g1", "g3"}); }));
    auto gc = make_grid_points_cartesian_product<2>(c, L, {"g1", "g3"});
    CHECK(gc.size() == 6 && gc.points[1][0] == 1 && gc.points[1][1] == 8 &&
          gc.points[2][0] == 2);
  }

  // The 12-bin x 15-radius Delta-Sigma configuration of the batched stream,
  // written as a module config.
  DataBlock
  dsigma_config(bool cartesian, std::string const& algorithm = "pagani")
  {
    DataBlock c;
    std::string const L = SigmaMiscent6DIntegrand::module_label();
    c.put_val(L, "algorithm", algorithm);
    c.put_val(L, "eps_rel", 1e-3);
    c.put_val(L, "eps_abs", 0);  // an int, as an ini parser may store it
    c.put_val(L, "max_eval", 100000000);
    c.put_val(L, "use_cartesian_product", cartesian);
    double const lnm_lo = std::log(1e13), lnm_hi = 15.5 * std::log(10.0);
    c.put_val(L, "lo_low", std::vector<double>{20, 30});
    c.put_val(L, "lo_high", std::vector<double>{30, 45});
    c.put_val(L, "lc_low", std::vector<double>{1, 1});
    c.put_val(L, "lc_high", std::vector<double>{300, 300});
    c.put_val(L, "zt_low", std::vector<double>{0.1, 0.25});
    c.put_val(L, "zt_high", std::vector<double>{0.45, 0.6});
    c.put_val(L, "lnm_low", std::vector<double>{lnm_lo, lnm_lo});
    c.put_val(L, "lnm_high", std::vector<double>{lnm_hi, lnm_hi});
    c.put_val(L, "rmis_low", std::vector<double>{0, 0});
    c.put_val(L, "rmis_high", std::vector<double>{1e30, 0.3});
    c.put_val(L, "theta_low", std::vector<double>{0, 0});
    c.put_val(L, "theta_high", std::vector<double>{2 * M_PI, 1.0});
    c.put_val(L, "zo_low", std::vector<double>{0.2, 0.35});
    c.put_val(L, "zo_high", std::vector<double>{0.35, 0.5});
    c.put_val(L, "radii", std::vector<double>{0.5, 2.0});
    return c;
  }

  void
  test_module_config()
  {
    using M = SigmaMiscent6DIntegrand;
    {
      DataBlock c = dsigma_config(false);
      auto m = parse_module_config<M>(c);
      CHECK(m.algorithm == Y3G_PAGANI && m.eps_rel == 1e-3 && m.eps_abs == 0.0);
      CHECK(m.max_eval == 1e8 && !m.cartesian && m.volumes.size() == 2 && m.grid.size() == 2);
      CHECK(m.num_results() == 2);
      CHECK(m.options.precision == Y3G_FP32SPLIT);
    }
    {
      DataBlock c = dsigma_config(true, "mcubes");
      auto m = parse_module_config<M>(c);
      CHECK(m.algorithm == Y3G_MCUBES && m.cartesian && m.num_results() == 4);
    }
    {
      DataBlock c = dsigma_config(false, "suave");
      CHECK(throws<std::runtime_error>([&] { parse_module_config<M>(c); },
                                       "unrecognized algorithm"));
    }
    {
      DataBlock c = dsigma_config(false);
      c.replace_val(M::module_label(), "radii", std::vector<double>{1, 2, 3});
      c.replace_val(M::module_label(), "zo_low", std::vector<double>{0.2, 0.35, 0.5});
      c.replace_val(M::module_label(), "zo_high", std::vector<double>{0.35, 0.5, 0.65});
      CHECK(throws<std::runtime_error>([&] { parse_module_config<M>(c); },
                                       "did not equal the number of gridpoints"));
    }
    {
      DataBlock c = dsigma_config(false);
      c.replace_val(M::module_label(), "max_eval", 0);
      CHECK(throws<std::runtime_error>([&] { parse_module_config<M>(c); }, "maxcalls"));
    }
    {
      DataBlock c = dsigma_config(false);
      c.put_val(M::module_label(), "precision", std::string("fp64"));
      CHECK(parse_module_config<M>(c).options.precision == Y3G_FP64);
      c.replace_val(M::module_label(), "precision", std::string("fp16"));
      CHECK(throws<std::runtime_error>([&] { parse_module_config<M>(c); }, "precision"));
    }
  }

  void
  test_integral_list()
  {
    using M = SigmaMiscent6DIntegrand;
    DataBlock c = dsigma_config(true);
    auto m = parse_module_config<M>(c);
    M ig(c);
    ig.set_spec(m.spec);
    DataBlock s;
    s.put_val("cosmological_parameters", "omega_m", 0.3);
    ig.set_sample(s);
    auto list = build_integrals(ig, m);
    CHECK(list.size() == 4);
    // volume-major: [v0 g0, v0 g1, v1 g0, v1 g1]
    CHECK(list[0].radius == 0.5 && list[1].radius == 2.0 && list[2].radius == 0.5);
    CHECK(list[1].lo_lo == 20 && list[2].lo_lo == 30);
    CHECK(list[1].zo_lo == 0.35 && list[1].zo_hi == 0.5);
    for (auto const& g : list) {
      CHECK(g.kind == Y3G_DSIGMA6D && g.ndim == 6 && g.algorithm == Y3G_PAGANI);
      CHECK(g.eps_rel == 1e-3 && g.eps_abs == 0.0);
    }
    // full circle folds onto [0, pi]; the offset radius maps to its
    // Rayleigh CDF, cut at 6 sigma
    y3g_params dp;
    y3g_default_params(&dp);
    double const sg = dp.mis_rlambda * std::pow(25.0 / 100.0, 0.2);
    CHECK(close(list[0].sigma_mis, sg));
    CHECK(list[0].lows[5] == 0.0 && close(list[0].highs[5], M_PI) && list[0].rmis_norm == 1.0);
    CHECK(list[0].lows[4] == 0.0 && close(list[0].highs[4], -std::expm1(-18.0)));
    CHECK(close(list[0].lnlt_lo, 0.0) && close(list[0].lnlt_hi, std::log(300.0)));
    CHECK(list[0].lows[2] == 0.1 && list[0].highs[2] == 0.45);
    // a partial angle keeps its range with half the weight
    double const sg1 = dp.mis_rlambda * std::pow(37.5 / 100.0, 0.2);
    CHECK(list[2].lows[5] == 0.0 && list[2].highs[5] == 1.0 && list[2].rmis_norm == 0.5);
    CHECK(close(list[2].highs[4], -std::expm1(-0.5 * (0.3 / sg1) * (0.3 / sg1))));
    // the standard descriptor of the library for the same integral
    y3g_table_spec sp;
    y3g_spec_wide(&sp);
    y3g_integral ref;
    y3g_make_integral(&dp, &sp, Y3G_DSIGMA6D, 20, 30, 0.2, 0.35, 0.5, &ref);
    for (int a = 0; a < 6; ++a) {
      if (!(close(list[0].lows[a], ref.lows[a]) && close(list[0].highs[a], ref.highs[a])))
        std::fprintf(stderr, "axis %d: [%.17g, %.17g] vs [%.17g, %.17g]\n", a,
                     list[0].lows[a], list[0].highs[a], ref.lows[a], ref.highs[a]);
      CHECK(close(list[0].lows[a], ref.lows[a]) && close(list[0].highs[a], ref.highs[a]));
    }
    // invalid volumes are reported, not integrated
    volume_t bad{{20, 1, 0.1, 30, 0, 0}, {30, 300, 1.5, 35, 1, 1}};  // zt beyond tables
    ig.set_grid_point({0.2, 0.35, 1.0});
    CHECK(throws<std::runtime_error>([&] { ig.describe(bad); }, "zt range"));
    volume_t bad2{{20, 1, 0.1, 30, 0, 0}, {30, 300, 0.4, 35, 1, 7.0}};
    CHECK(throws<std::runtime_error>([&] { ig.describe(bad2); }, "theta"));
  }

  void
  test_counts_lists()
  {
    DataBlock c;
    std::string const L = NumberCounts5DIntegrand::module_label();
    c.put_val(L, "algorithm", std::string("pagani"));
    c.put_val(L, "eps_rel", 1e-4);
    c.put_val(L, "eps_abs", 1e-12);
    c.put_val(L, "max_eval", 1e8);
    c.put_val(L, "use_cartesian_product", false);
    for (char const* n : {"lo", "zo", "lc", "zt", "lnm"}) {
      c.put_val(L, std::string(n) + "_low", std::vector<double>{1, 2, 3});
      c.put_val(L, std::string(n) + "_high", std::vector<double>{2, 3, 4});
    }
    c.replace_val(L, "zt_low", std::vector<double>{0.1, 0.1, 0.1});
    c.replace_val(L, "zt_high", std::vector<double>{0.5, 0.5, 0.5});
    auto m = parse_module_config<NumberCounts5DIntegrand>(c);
    CHECK(m.volumes.size() == 3 && m.num_results() == 3);
    NumberCounts5DIntegrand ig(c);
    ig.set_spec(m.spec);
    DataBlock s;
    s.put_val("cosmological_parameters", "omega_m", 0.25);
    s.put_val("y3g_synthetic", "f_prj", 0.1);
    ig.set_sample(s);
    CHECK(ig.params().omega_m == 0.25 && ig.params().f_prj == 0.1);
    auto list = build_integrals(ig, m);
    CHECK(list.size() == 3 && list[2].kind == Y3G_COUNT5D && list[2].ndim == 5);
    // z_obs is integrated as its standardised photo-z variable: the bin goes
    // to zo_lo / zo_hi, the box axis is [0, 1]
    CHECK(list[2].lo_lo == 3 && list[2].zo_lo == 3 && list[2].zo_hi == 4 && list[2].lows[0] == 3 &&
          list[2].lows[1] == 0 && list[2].highs[1] == 1);
    CHECK(close(list[1].lnlt_lo, std::log(2.0)) && close(list[1].lnlt_hi, std::log(3.0)));
    CHECK(list[1].lows[4] == 2 && list[1].highs[4] == 3);
    // the cartesian layout of a grid-less module is [nvol, 1]
    c.replace_val(L, "use_cartesian_product", true);
    auto mc = parse_module_config<NumberCounts5DIntegrand>(c);
    CHECK(mc.num_results() == 3 && mc.num_grid_points() == 1);
    std::vector<y3g_result> r(3);
    DataBlock out;
    put_results(out, mc, r);
    auto gp = out.view<ndarray<double>>(L, "gridpoints");
    CHECK(gp.extents() == (std::vector<std::size_t>{1, 0}));
  }

  void
  test_put_results()
  {
    using M = SigmaMiscent6DIntegrand;
    std::vector<y3g_result> r(4);
    for (int i = 0; i < 4; ++i) {
      r[i].value = 10.0 + i;
      r[i].error = 0.1 * i;
      r[i].prob = 0.5 + i;
      r[i].neval = 100 * i;
      r[i].nregions = 7 * i;
      r[i].status = i % 2;
    }
    {
      DataBlock c = dsigma_config(true);
      auto m = parse_module_config<M>(c);
      DataBlock s;
      put_results(s, m, r);
      std::string const L = M::module_label();
      auto vals = s.view<ndarray<double>>(L, "vals");
      CHECK(vals.extents() == (std::vector<std::size_t>{2, 2}));
      CHECK(vals(1, 0) == 12.0 && vals(0, 1) == 11.0);
      CHECK(s.view<ndarray<double>>(L, "probs")(1, 1) == 3.5);  // chi^2 kept
      CHECK(s.view<ndarray<int>>(L, "status")(0, 1) == 1);
      CHECK(s.view<ndarray<int>>(L, "nregions")(1, 1) == 21);
      auto ivol = s.view<ndarray<int>>(L, "ivol");
      CHECK(ivol(0, 1) == 0 && ivol(1, 0) == 1);
      auto gp = s.view<ndarray<double>>(L, "gridpoints");
      CHECK(gp.extents() == (std::vector<std::size_t>{2, 3}));
      CHECK(gp(1, 0) == 0.35 && gp(1, 2) == 2.0);
      CHECK(s.view<ndarray<double>>(L, "errors")(1, 1) == 0.1 * 3);
    }
    {
      DataBlock c = dsigma_config(false);
      auto m = parse_module_config<M>(c);
      DataBlock s;
      std::vector<y3g_result> r2(r.begin(), r.begin() + 2);
      put_results(s, m, r2);
      std::string const L = M::module_label();
      auto probs = s.view<ndarray<double>>(L, "probs");
      CHECK(probs.extents() == (std::vector<std::size_t>{2}) && probs(1) == 0.0);
      CHECK(!s.has_val(L, "ivol") && !s.has_val(L, "gridpoints"));
      std::vector<double> v;
      CHECK(s.get_val(L, "vals", v) == DBS_SUCCESS && v == (std::vector<double>{10, 11}));
      // a second put into the same sample is an error, as in CosmoSIS
      CHECK(throws<std::runtime_error>([&] { put_results(s, m, r2); }));
    }
  }

  // The reference's table sections in a sample (mass_function, distances).
  void
  test_table_inputs()
  {
    DataBlock cfg;
    NumberCounts5DIntegrand ig(cfg);
    y3g_table_spec sp;
    y3g_spec_wide(&sp);
    ig.set_spec(sp);
    auto sample = [](int nm, bool uniform, bool good_shape) {
      DataBlock s;
      s.put_val("cosmological_parameters", "omega_M", 0.3);  // = omega_m
      s.put_val("cosmological_parameters", "omega_nu", 0.0);
      s.put_val("cluster_abundance", "hmf_s", 0.1);
      s.put_val("cluster_abundance", "hmf_q", 0.9);
      std::vector<double> mh(nm), z(8), vals;
      for (int i = 0; i < nm; ++i) {
        double const lnm = 30.0 + 5.0 * i / (nm - 1) + (uniform || i != 3 ? 0.0 : 0.01);
        mh[i] = std::exp(lnm) / 0.3;
      }
      for (int j = 0; j < 8; ++j) z[j] = 0.1 + 0.1 * j;
      for (int j = 0; j < 8; ++j)
        for (int i = 0; i < nm; ++i) vals.push_back(100.0 * j + i);
      s.put_val("mass_function", "m_h", mh);
      s.put_val("mass_function", "z", z);
      if (good_shape)
        s.put_val("mass_function", "dndlnmh", ndarray<double>(vals, {8, std::size_t(nm)}));
      else
        s.put_val("mass_function", "dndlnmh", ndarray<double>(vals, {std::size_t(nm), 8}));
      std::vector<double> dz(101), da(101);
      for (int i = 0; i < 101; ++i) dz[i] = 0.01 * i, da[i] = 3000.0 * dz[i];
      s.put_val("distances", "z", dz);
      s.put_val("distances", "d_a", da);
      return s;
    };
    DataBlock s = sample(50, true, true);
    ig.set_sample(s);
    auto const& q = ig.spec();
    CHECK(q.n_lnm == 50 && q.n_z == 8 && close(q.lnm_lo, 30.0, 1e-12) &&
          close(q.lnm_hi, 35.0, 1e-12) && close(q.z_lo, 0.1) && close(q.z_hi, 0.8));
    CHECK(q.n_da == 101 && close(q.da_zmax, 1.0));
    CHECK(ig.hmf_table().size() == 400 && ig.hmf_table()[51] == 101.0);
    CHECK(ig.da_table().size() == 101 && ig.da_table()[100] == 3000.0);
    CHECK(ig.params().omega_m == 0.3 && ig.params().hmf_s == 0.1 && ig.params().hmf_q == 0.9);
    DataBlock bad = sample(50, false, true);
    CHECK(throws<std::runtime_error>([&] { ig.set_sample(bad); }, "not uniformly spaced"));
    DataBlock bad2 = sample(50, true, false);
    CHECK(throws<std::runtime_error>([&] { ig.set_sample(bad2); }, "shape"));
    // without the sections the configured spec comes back
    DataBlock plain;
    plain.put_val("cosmological_parameters", "omega_m", 0.3);
    ig.set_sample(plain);
    CHECK(ig.spec().n_lnm == sp.n_lnm && ig.hmf_table().empty() && ig.da_table().empty());
  }

  // setup / execute / cleanup through dlopen, as CosmoSIS calls them.
  void
  test_entry_points(char const* module_path, bool have_gpu)
  {
    void* h = dlopen(module_path, RTLD_NOW | RTLD_LOCAL);
    CHECK(h != nullptr);
    if (!h) {
      std::fprintf(stderr, "dlopen: %s\n", dlerror());
      return;
    }
    using setup_t = void* (*)(cosmosis::DataBlock*);
    using execute_t = DATABLOCK_STATUS (*)(cosmosis::DataBlock*, void*);
    using cleanup_t = int (*)(void*);
    auto setup = reinterpret_cast<setup_t>(dlsym(h, "setup"));
    auto execute = reinterpret_cast<execute_t>(dlsym(h, "execute"));
    auto cleanup = reinterpret_cast<cleanup_t>(dlsym(h, "cleanup"));
    CHECK(setup && execute && cleanup);
    if (!(setup && execute && cleanup)) return;
    DataBlock bad = dsigma_config(false, "vegas");
    CHECK(setup(&bad) == nullptr);  // unknown algorithm: no object, no throw
    CHECK(setup(nullptr) == nullptr);
    DataBlock sample;
    CHECK(execute(&sample, nullptr) != DBS_SUCCESS);
    CHECK(execute(nullptr, nullptr) != DBS_SUCCESS);
    CHECK(cleanup(nullptr) == 0);
    DataBlock good = dsigma_config(false);
    void* mod = setup(&good);
    if (!have_gpu) {
      CHECK(mod == nullptr);  // no device to bind
    } else {
      CHECK(mod != nullptr);
      DataBlock empty;  // no cosmological_parameters: an error status
      CHECK(execute(&empty, mod) != DBS_SUCCESS);
      CHECK(cleanup(mod) == 0);
    }
    dlclose(h);
  }
} // namespace

int
main(int argc, char** argv)
{
  test_datablock();
  test_readers();
  test_volumes_and_grids();
  test_module_config();
  test_integral_list();
  test_counts_lists();
  test_put_results();
  test_table_inputs();
  if (argc > 1) {
    y3g_runtime* rt = nullptr;
    bool const have_gpu = y3g_runtime_create(0, &rt) == Y3G_OK;
    if (rt) y3g_runtime_destroy(rt);
    test_entry_points(argv[1], have_gpu);
  }
  std::printf("%d checks, %d failures\n", checks, failures);
  return failures == 0 ? 0 : 1;
}
