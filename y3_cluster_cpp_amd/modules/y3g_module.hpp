// y3g_module.hpp -- host side of the CosmoSIS integration modules built on
// liby3g (include/y3g.h).
//
// A module is one shared library exporting the three CosmoSIS entry points
// setup / execute / cleanup (reference: src/utils/cuda_module_macros.cuh:20-45)
// around HIPIntegrationModule<I>, the counterpart of the reference's
// CosmoSISSICUDAModule<I> (src/utils/CosmoSISSICUDAModule.hh:99-467):
//
//   setup    reads <label>/{algorithm, eps_rel, eps_abs, max_eval,
//            use_cartesian_product} and the integration volumes and grid
//            points, and binds the GPU of the process's node-local rank;
//   execute  reads the sample's parameters, (re)builds the per-sample tables
//            on the device when they changed, integrates every
//            (volume, grid point) pair in ONE device submission, and puts
//            vals / errors / probs / status / nregions (+ ivol, gridpoints
//            for the cartesian product) into the sample's <label> section;
//   cleanup  deletes the module.
//
// Unlike the reference, exceptions never cross the extern "C" edge: setup
// returns nullptr and execute a non-zero DATABLOCK_STATUS after printing the
// error.  Every helper here is host code; the device work happens inside
// liby3g.
#ifndef Y3G_MODULE_HPP
#define Y3G_MODULE_HPP

#include "cosmosis/datablock/datablock.hh"
#include "cosmosis/datablock/ndarray.hh"
#include "y3g.h"

#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace y3g_cosmosis {

  // Optional parts of a module's integrand class: upload(y3g_runtime*) for
  // per-sample device data beyond the tables, and a static fp64_only for
  // integrands that have no FP32 path.
  template <class I, class = void>
  struct has_device_upload : std::false_type {};
  template <class I>
  struct has_device_upload<
    I, std::void_t<decltype(std::declval<I&>().upload(std::declval<y3g_runtime*>()))>>
    : std::true_type {};
  template <class I, class = void>
  struct is_fp64_only : std::false_type {};
  template <class I>
  struct is_fp64_only<I, std::void_t<decltype(I::fp64_only)>>
    : std::bool_constant<I::fp64_only> {};

  // ---------------------------------------------------------- config readers

  // A std::vector<double> stored as vector<double>, double, vector<int> or
  // int (reference: src/utils/datablock_reader.cc:3-28, same error text).
  inline std::vector<double>
  get_vector_double(cosmosis::DataBlock& db, std::string const& section,
                    std::string const& name)
  {
    std::vector<double> v;
    if (db.get_val(section, name, v) == DBS_SUCCESS) return v;
    double x = 0.0;
    if (db.get_val(section, name, x) == DBS_SUCCESS) return {x};
    std::vector<int> iv;
    if (db.get_val(section, name, iv) == DBS_SUCCESS)
      return std::vector<double>(iv.begin(), iv.end());
    int i = 0;
    if (db.get_val(section, name, i) == DBS_SUCCESS) return {double(i)};
    throw std::runtime_error("DataBlock contained no parameer in section " +
                             section + " named " + name +
                             " that could be coerced to type vector<double>");
  }

  // A scalar stored as double or int (reference:
  // src/utils/CosmoSISSICUDAModule.hh:37-50).
  inline double
  view_as_double(cosmosis::DataBlock& cfg, std::string const& section,
                 std::string const& name)
  {
    double x = 0.0;
    if (cfg.get_val(section, name, x) == DBS_SUCCESS) return x;
    int i = 0;
    if (cfg.get_val(section, name, i) == DBS_SUCCESS) return double(i);
    throw std::runtime_error("CosmoSISSICUDAModule: parameter " + section + "/" +
                             name + " could not be read as double or int");
  }

  inline double
  optional_double(cosmosis::DataBlock& db, std::string const& section,
                  std::string const& name, double def)
  {
    return db.has_val(section, name) ? view_as_double(db, section, name) : def;
  }

  // ------------------------------------------------------ volumes and grids

  struct volume_t {
    std::vector<double> lows, highs;
  };

  // One volume per entry of <names[0]>_low; every <name>_low / _high list
  // has that length (reference:
  // src/utils/make_cuda_integration_volumes.cuh:48-91, 168-194).
  inline std::vector<volume_t>
  make_integration_volumes_wall_of_numbers(cosmosis::DataBlock& cfg,
                                           std::string const& label,
                                           std::vector<std::string> const& names)
  {
    std::vector<volume_t> out;
    if (names.empty()) return out;
    std::size_t nvol = 0;
    for (std::size_t a = 0; a != names.size(); ++a) {
      auto lo = get_vector_double(cfg, label, names[a] + "_low");
      auto hi = get_vector_double(cfg, label, names[a] + "_high");
      if (a == 0) {
        nvol = lo.size();
        out.assign(nvol, volume_t{std::vector<double>(names.size()),
                                  std::vector<double>(names.size())});
      }
      if (lo.size() != nvol || hi.size() != nvol)
        throw std::runtime_error(names[a] + " bad, bad user!");
      for (std::size_t v = 0; v != nvol; ++v) {
        out[v].lows[a] = lo[v];
        out[v].highs[a] = hi[v];
      }
    }
    return out;
  }

  // Odometer over lists, first list slowest (reference:
  // src/utils/meta.hh:24-34).
  inline std::vector<std::vector<double>>
  cartesian_product(std::vector<std::vector<double>> const& axes)
  {
    std::vector<std::vector<double>> out;
    if (axes.empty()) return out;
    for (auto const& a : axes)
      if (a.empty()) return out;
    std::vector<std::size_t> k(axes.size(), 0);
    while (true) {
      std::vector<double> pt(axes.size());
      for (std::size_t i = 0; i != axes.size(); ++i) pt[i] = axes[i][k[i]];
      out.push_back(pt);
      std::size_t i = axes.size();
      while (i > 0) {
        --i;
        if (++k[i] < axes[i].size()) break;
        k[i] = 0;
        if (i == 0) return out;
      }
    }
  }

  // Lows and highs each taken as a cartesian product of the per-variable
  // lists (reference: src/utils/make_cuda_integration_volumes.cuh:157-183,
  // 196-222).
  inline std::vector<volume_t>
  make_integration_volumes_cartesian_product(cosmosis::DataBlock& cfg,
                                             std::string const& label,
                                             std::vector<std::string> const& names)
  {
    std::vector<std::vector<double>> lo, hi;
    for (auto const& n : names) {
      lo.push_back(get_vector_double(cfg, label, n + "_low"));
      hi.push_back(get_vector_double(cfg, label, n + "_high"));
    }
    auto L = cartesian_product(lo), H = cartesian_product(hi);
    if (L.size() != H.size())
      throw std::runtime_error("cartesian product of volume bounds: lows and highs "
                               "differ in number");
    std::vector<volume_t> out(L.size());
    for (std::size_t i = 0; i != L.size(); ++i) out[i] = {L[i], H[i]};
    return out;
  }

  // Points of a space with NAXES named axes (reference:
  // src/utils/make_grid_points.hh:18-60).
  template <std::size_t NAXES>
  struct grid_t {
    using value_type = std::array<double, NAXES>;
    std::vector<value_type> points;
    std::vector<std::string> names;
    std::size_t size() const { return points.size(); }
    static constexpr std::size_t naxes() { return NAXES; }
  };

  // Axis lists zipped into points; unequal lengths are an error (reference:
  // src/utils/make_grid_points.hh:224-240, 267-284).
  template <std::size_t N>
  grid_t<N>
  make_grid_points_wall_of_numbers(cosmosis::DataBlock& cfg,
                                   std::string const& label,
                                   std::array<std::string, N> const& names)
  {
    grid_t<N> g;
    g.names.assign(names.begin(), names.end());
    if constexpr (N == 0) {
      return g;
    } else {
      std::array<std::vector<double>, N> axes;
      for (std::size_t i = 0; i != N; ++i)
        axes[i] = get_vector_double(cfg, label, names[i]);
      for (std::size_t i = 1; i != N; ++i)
        if (axes[i].size() != axes[0].size())
          throw std::invalid_argument("unequal length axes in make_grid_wall_of_numbers");
      g.points.resize(axes[0].size());
      for (std::size_t p = 0; p != axes[0].size(); ++p)
        for (std::size_t i = 0; i != N; ++i) g.points[p][i] = axes[i][p];
      return g;
    }
  }

  // Cartesian product of the axis lists, first axis slowest (reference:
  // src/utils/make_grid_points.hh:212-220, 247-265).
  template <std::size_t N>
  grid_t<N>
  make_grid_points_cartesian_product(cosmosis::DataBlock& cfg,
                                     std::string const& label,
                                     std::array<std::string, N> const& names)
  {
    grid_t<N> g;
    g.names.assign(names.begin(), names.end());
    std::vector<std::vector<double>> axes;
    for (auto const& n : names) axes.push_back(get_vector_double(cfg, label, n));
    for (auto const& pt : cartesian_product(axes)) {
      typename grid_t<N>::value_type v{};
      for (std::size_t i = 0; i != N; ++i) v[i] = pt[i];
      g.points.push_back(v);
    }
    return g;
  }

  // -------------------------------------------------------- process binding

  // Node-local and global rank from the launcher's environment (the
  // reference asks MPI, src/utils/mpi_support.cc:11-33; the variables below
  // are what Open MPI, MPICH/PMI, Slurm and torchrun export).
  struct rank_info {
    int local_rank = 0;
    int global_rank = 0;
  };

  inline int
  env_int(std::initializer_list<char const*> names, int def)
  {
    for (char const* n : names)
      if (char const* v = std::getenv(n)) return std::atoi(v);
    return def;
  }

  inline rank_info
  get_rank_info()
  {
    rank_info r;
    r.local_rank = env_int({"OMPI_COMM_WORLD_LOCAL_RANK", "MPI_LOCALRANKID",
                            "SLURM_LOCALID", "LOCAL_RANK"}, 0);
    r.global_rank = env_int({"OMPI_COMM_WORLD_RANK", "PMI_RANK", "SLURM_PROCID",
                             "RANK"}, 0);
    return r;
  }

  inline void
  record_timestamp(std::ostream& os, char const* label)
  {
    using namespace std::chrono;
    double const t =
      duration<double>(system_clock::now().time_since_epoch()).count();
    os << label << '\t' << std::setprecision(17) << t << std::endl;
  }

  // One device binding per module object: the GPU whose index is the
  // node-local rank (reference DeviceInitializer,
  // src/utils/CosmoSISSICUDAModule.hh:59-83).
  class DeviceBinding {
  public:
    explicit DeviceBinding(int device) : id(device)
    {
      if (y3g_runtime_create(device, &rt_) != Y3G_OK)
        throw std::runtime_error(std::string("Failed to bind GPU ") +
                                 std::to_string(device) + ": " + y3g_last_error());
    }
    ~DeviceBinding()
    {
      if (rt_) y3g_runtime_destroy(rt_);
    }
    DeviceBinding(DeviceBinding const&) = delete;
    DeviceBinding& operator=(DeviceBinding const&) = delete;
    y3g_runtime* get() const { return rt_; }
    int id;

  private:
    y3g_runtime* rt_ = nullptr;
  };

  // ------------------------------------------------------- module settings

  // Everything setup() reads, parsed before any device is touched.
  template <class I>
  struct module_config {
    int algorithm = Y3G_PAGANI;
    double eps_rel = 0.0, eps_abs = 0.0, max_eval = 0.0;
    bool cartesian = false;
    bool do_timing = false;
    std::string memory_track_file;  // empty: no memory log
    std::vector<volume_t> volumes;
    typename I::grid_t grid;
    y3g_options options{};
    y3g_table_spec spec{};

    // Number of integrals of one execute(): volumes x grid points for the
    // cartesian product, else the zipped count.
    std::size_t
    num_results() const
    {
      std::size_t const ng = I::grid_t::naxes() == 0 ? 1 : grid.size();
      return cartesian ? volumes.size() * ng : volumes.size();
    }
    std::size_t
    num_grid_points() const
    {
      return I::grid_t::naxes() == 0 ? 1 : grid.size();
    }
  };

  // "pagani" / "mcubes", anything else an error (reference:
  // src/utils/gpu_integrator.cuh:54-65).
  inline int
  algorithm_from_name(std::string const& name)
  {
    if (name == "pagani") return Y3G_PAGANI;
    if (name == "mcubes") return Y3G_MCUBES;
    throw std::runtime_error("MultiDimensionalIntegrator::integrate called for "
                             "unrecognized algorithm");
  }

  template <class I>
  module_config<I>
  parse_module_config(cosmosis::DataBlock& cfg)
  {
    std::string const L = I::module_label();
    module_config<I> m;
    m.algorithm = algorithm_from_name(cfg.view<std::string>(L, "algorithm"));
    m.volumes = I::make_integration_volumes(cfg);
    m.grid = I::make_grid_points(cfg);
    m.eps_rel = view_as_double(cfg, L, "eps_rel");
    m.eps_abs = view_as_double(cfg, L, "eps_abs");
    m.cartesian = cfg.view<bool>(L, "use_cartesian_product");
    m.max_eval = view_as_double(cfg, L, "max_eval");
    if (m.max_eval <= 0.0)
      throw std::runtime_error("MultiDimensionalIntegrator: maxcalls can not be negative\n");
    if (!m.cartesian && I::grid_t::naxes() != 0 && m.volumes.size() != m.grid.size())
      throw std::runtime_error(
        "An integration module was configured to use a zipped sequence of "
        "volumes and gridpoints,\nbut the number of volumes did not equal the "
        "number of gridpoints.\n");
    bool timing = false;
    if (cfg.get_val(L, "do_timing", timing) == DBS_SUCCESS) m.do_timing = timing;
    std::string mem;
    if (cfg.get_val(L, "memory_track_file", mem) == DBS_SUCCESS) m.memory_track_file = mem;

    // device-side settings (defaults from the library)
    y3g_options_default(&m.options);
    std::string prec = is_fp64_only<I>::value ? "fp64" : "fp32split";
    cfg.get_val(L, "precision", prec);
    if (prec == "fp64")
      m.options.precision = Y3G_FP64;
    else if (prec == "fp32split" && !is_fp64_only<I>::value)
      m.options.precision = Y3G_FP32SPLIT;
    else
      throw std::runtime_error(L + "/precision must be " +
                               (is_fp64_only<I>::value ? "fp64" : "fp32split or fp64") +
                               ", not " + prec);
    m.options.max_regions =
      int(optional_double(cfg, L, "max_regions", m.options.max_regions));
    m.options.vegas_calls =
      (long long)optional_double(cfg, L, "vegas_calls", double(m.options.vegas_calls));
    m.options.vegas_itmax = int(optional_double(cfg, L, "vegas_itmax", m.options.vegas_itmax));
    // the synthetic integrands' HMF: hat-smoothed (1, default) or bilinear (0)
    m.options.hmf_smooth = optional_double(cfg, L, "hmf_smooth", m.options.hmf_smooth) != 0.0;
    // tolerance-band budgets (the autotuner's bands): ascending eps edges
    if (cfg.has_val(L, "band_eps") || cfg.has_val(L, "band_regions")) {
      auto const be = get_vector_double(cfg, L, "band_eps");
      auto const br = get_vector_double(cfg, L, "band_regions");
      if (be.size() != br.size() || be.size() > 8)
        throw std::runtime_error(L + "/band_eps and band_regions must have the same length "
                                     "(at most 8)");
      m.options.n_bands = int(be.size());
      for (std::size_t k = 0; k < be.size(); ++k) {
        m.options.band_eps[k] = be[k];
        m.options.band_regions[k] = int(br[k]);
        if (m.options.band_regions[k] > m.options.max_regions)
          m.options.max_regions = m.options.band_regions[k];
      }
    }

    y3g_spec_wide(&m.spec);
    m.spec.lnm_lo = optional_double(cfg, L, "table_lnm_low", m.spec.lnm_lo);
    m.spec.lnm_hi = optional_double(cfg, L, "table_lnm_high", m.spec.lnm_hi);
    m.spec.z_lo = optional_double(cfg, L, "table_z_low", m.spec.z_lo);
    m.spec.z_hi = optional_double(cfg, L, "table_z_high", m.spec.z_hi);
    return m;
  }

  // ------------------------------------------------------------- results

  // The integral list of one execute(): volume-major over the grid for the
  // cartesian product (reference: CosmoSISSICUDAModule.hh:282-313), else
  // volume i with grid point i (:315-335).
  template <class I>
  std::vector<y3g_integral>
  build_integrals(I& integrand, module_config<I> const& m)
  {
    std::vector<y3g_integral> out;
    out.reserve(m.num_results());
    auto add = [&](volume_t const& v) {
      y3g_integral ig = integrand.describe(v);
      ig.algorithm = m.algorithm;
      ig.eps_rel = m.eps_rel;
      ig.eps_abs = m.eps_abs;
      // PAGANI takes no call limit (reference gpu_integrator.cuh:170); the
      // region budget bounds it
      ig.max_eval = m.algorithm == Y3G_MCUBES ? (long long)m.max_eval : (1LL << 62);
      ig.point = 0;
      out.push_back(ig);
    };
    std::size_t const ng = m.num_grid_points();
    if (m.cartesian) {
      for (auto const& v : m.volumes)
        for (std::size_t g = 0; g != ng; ++g) {
          if constexpr (I::grid_t::naxes() != 0) integrand.set_grid_point(m.grid.points[g]);
          add(v);
        }
    } else {
      for (std::size_t i = 0; i != m.volumes.size(); ++i) {
        if constexpr (I::grid_t::naxes() != 0) integrand.set_grid_point(m.grid.points[i]);
        add(m.volumes[i]);
      }
    }
    return out;
  }

  // Put the result arrays into the sample (reference:
  // CosmoSISSICUDAModule.hh:350-467).  probs is the chi^2 of m-Cubes in
  // the cartesian layout and 0.0 in the zipped one, as the reference stores
  // them (:306, :329).
  template <class I>
  void
  put_results(cosmosis::DataBlock& sample, module_config<I> const& m,
              std::vector<y3g_result> const& res)
  {
    using cosmosis::ndarray;
    std::string const L = I::module_label();
    std::size_t const n = m.num_results();
    if (res.size() != n) throw std::logic_error("put_results: result count");
    std::vector<double> vals(n), errs(n), probs(n);
    std::vector<int> stat(n), nreg(n);
    for (std::size_t i = 0; i != n; ++i) {
      vals[i] = res[i].value;
      errs[i] = res[i].error;
      probs[i] = m.cartesian ? res[i].prob : 0.0;
      stat[i] = res[i].status;
      nreg[i] = res[i].nregions;
    }
    std::vector<std::size_t> ext;
    if (m.cartesian)
      ext = {m.volumes.size(), m.num_grid_points()};
    else
      ext = {n};
    auto put = [&](char const* key, auto const& arr) {
      if (sample.put_val(L, key, arr) != DBS_SUCCESS)
        throw std::runtime_error(L + "/" + key + " could not be put into the sample");
    };
    put("vals", ndarray<double>(vals, ext));
    put("errors", ndarray<double>(errs, ext));
    put("probs", ndarray<double>(probs, ext));
    put("status", ndarray<int>(stat, ext));
    put("nregions", ndarray<int>(nreg, ext));
    if (m.cartesian) {
      std::size_t const ng = m.num_grid_points(), na = I::grid_t::naxes();
      std::vector<int> ivol(n);
      for (std::size_t v = 0; v != m.volumes.size(); ++v)
        for (std::size_t g = 0; g != ng; ++g) ivol[v * ng + g] = int(v);
      std::vector<double> gp(ng * na);
      if constexpr (I::grid_t::naxes() != 0)
        for (std::size_t g = 0; g != ng; ++g)
          for (std::size_t a = 0; a != na; ++a) gp[g * na + a] = m.grid.points[g][a];
      put("ivol", ndarray<int>(ivol, ext));
      put("gridpoints", ndarray<double>(gp, {ng, na}));
    }
  }

  // ------------------------------------------------------------- the module

  // The device tables of parameter point 0: built from the sample's physics
  // and overwritten by the sample's own mass_function / distances tables
  // when it carries them; rebuilt only when any of those inputs changed.
  class TableCache {
  public:
    template <class S>
    void
    ensure(y3g_runtime* rt, S const& smp)
    {
      y3g_params const& p = smp.params();
      y3g_table_spec const& sp = smp.spec();
      if (have_ && std::memcmp(&p, &built_, sizeof(p)) == 0 &&
          std::memcmp(&sp, &built_spec_, sizeof(sp)) == 0 &&
          smp.hmf_table() == built_hmf_ && smp.da_table() == built_da_)
        return;
      have_ = false;
      if (y3g_tables_build(rt, &p, &sp, 1) != Y3G_OK)
        throw std::runtime_error(std::string("table build: ") + y3g_last_error());
      auto const& h = smp.hmf_table();
      auto const& d = smp.da_table();
      if ((!h.empty() || !d.empty()) &&
          y3g_tables_upload(rt, 0, h.empty() ? nullptr : h.data(),
                            d.empty() ? nullptr : d.data(), nullptr) != Y3G_OK)
        throw std::runtime_error(std::string("table upload: ") + y3g_last_error());
      built_ = p;
      built_spec_ = sp;
      built_hmf_ = h;
      built_da_ = d;
      have_ = true;
    }

  private:
    y3g_params built_{};
    y3g_table_spec built_spec_{};
    std::vector<double> built_hmf_, built_da_;
    bool have_ = false;
  };

  template <class I>
  class HIPIntegrationModule {
  public:
    explicit HIPIntegrationModule(cosmosis::DataBlock& cfg)
      : cfg_(parse_module_config<I>(cfg)),
        rank_(get_rank_info()),
        device_(rank_.local_rank),
        integrand_(cfg)
    {
      integrand_.set_spec(cfg_.spec);
      if (cfg_.do_timing) {
        std::ostringstream fn;
        fn << "mpi_timing_" << I::module_label() << '_' << rank_.global_rank << ".txt";
        timing_.emplace(fn.str());
        record_timestamp(*timing_, "constructed");
      }
      // memory log (CosmoSISSICUDAModule.hh:214-224): appended to, header
      // first; rows in MemTrackSentry's format (mem_tracking_sentry.hh:37-54)
      if (!cfg_.memory_track_file.empty()) {
        mem_track_.emplace(cfg_.memory_track_file, std::ios_base::app);
        *mem_track_ << "free_mem, module_mem, module_label, track_label, loc, "
                       "volume_id, grid_point_id"
                    << std::endl;
        track_memory("setup", "end");
      }
    }

    ~HIPIntegrationModule()
    {
      if (timing_) record_timestamp(*timing_, "destroyed");
    }

    void
    execute(cosmosis::DataBlock& sample)
    {
      if (timing_) record_timestamp(*timing_, "enter");
      track_memory("execute", "start");
      integrand_.set_sample(sample);
      tables_.ensure(device_.get(), integrand_);
      if constexpr (has_device_upload<I>::value) integrand_.upload(device_.get());
      std::vector<y3g_integral> igs = build_integrals(integrand_, cfg_);
      std::vector<y3g_result> res(igs.size());
      if (!igs.empty() &&
          y3g_integrate(device_.get(), igs.data(), igs.size(), &cfg_.options,
                        res.data()) != Y3G_OK)
        throw std::runtime_error(std::string("integrate: ") + y3g_last_error());
      put_results(sample, cfg_, res);
      track_memory("execute", "end");
      if (timing_) record_timestamp(*timing_, "exit");
    }

    module_config<I> const& config() const { return cfg_; }

  private:
    module_config<I> cfg_;
    rank_info rank_;
    DeviceBinding device_;
    I integrand_;
    TableCache tables_;
    std::optional<std::ofstream> timing_;
    std::optional<std::ofstream> mem_track_;

    // One row of the memory log: the device's free bytes, the bytes this
    // module's runtime holds, labels, and volume / grid point ids (-1: the
    // whole submission).
    void
    track_memory(char const* label, char const* loc)
    {
      if (!mem_track_) return;
      size_t fr = 0, tot = 0, mine = 0;
      if (y3g_runtime_memory(device_.get(), &fr, &tot, &mine) != Y3G_OK)
        std::cerr << "y3g_runtime_memory: " << y3g_last_error() << '\n';
      *mem_track_ << fr << "," << mine << "," << I::module_label() << "," << label << ","
                  << loc << "," << -1 << "," << -1 << std::endl;
    }
  };

  // ------------------------------------------------ the integrands' common part

  // Per-sample inputs of the integrands.
  //   physics   omega_m (and h0) from cosmological_parameters, the rest from
  //             the optional y3g_synthetic section (defaults
  //             y3g_default_params); hmf_s / hmf_q from cluster_abundance
  //             when present (HMF_t, hmf_t.cuh:44-52);
  //   tables    built on the device from the physics, unless the sample
  //             carries the reference's table sections:
  //               mass_function/{m_h, z, dndlnmh} -> the HMF table on the
  //                 axis ln(m_h (omega_M - omega_nu)) (hmf_t.cuh:10-20, 44-52);
  //               distances/{z, d_a} -> the D_A table (dv_do_dz_t.cuh:13-17).
  //             Their grids must be uniform (z from 0 for distances); the
  //             values are loaded as they are.
  class SyntheticSample {
  public:
    void
    set_sample(cosmosis::DataBlock& sample)
    {
      y3g_default_params(&p_);
      std::string const C = "cosmological_parameters", S = "y3g_synthetic";
      p_.omega_m = view_as_double(sample, C, "omega_m");  // = omega_M (case-insensitive)
      p_.h = optional_double(sample, C, "h0", p_.h);
      struct {
        char const* key;
        double* dst;
      } const keys[] = {
        {"hmf_amp", &p_.hmf_amp},         {"hmf_A", &p_.hmf_A},
        {"hmf_a", &p_.hmf_a},             {"hmf_c", &p_.hmf_c},
        {"mor_lnl0", &p_.mor_lnl0},       {"mor_slope", &p_.mor_slope},
        {"mor_zslope", &p_.mor_zslope},   {"mor_sigma_ln", &p_.mor_sigma_ln},
        {"f_prj", &p_.f_prj},             {"prj_scale", &p_.prj_scale},
        {"f_mis", &p_.f_mis},             {"mis_rlambda", &p_.mis_rlambda},
        {"z_source", &p_.z_source},       {"sigz0", &p_.sigz0}};
      for (auto const& k : keys) *k.dst = optional_double(sample, S, k.key, *k.dst);
      p_.hmf_s = optional_double(sample, "cluster_abundance", "hmf_s", p_.hmf_s);
      p_.hmf_q = optional_double(sample, "cluster_abundance", "hmf_q", p_.hmf_q);
      spec_ = base_spec_;
      hmf_tab_.clear();
      da_tab_.clear();
      if (sample.has_val("mass_function", "dndlnmh")) read_hmf_table_(sample);
      if (sample.has_val("distances", "d_a")) read_da_table_(sample);
    }
    void set_spec(y3g_table_spec const& s) { base_spec_ = spec_ = s; }
    y3g_params const& params() const { return p_; }
    y3g_table_spec const& spec() const { return spec_; }
    std::vector<double> const& hmf_table() const { return hmf_tab_; }
    std::vector<double> const& da_table() const { return da_tab_; }

  protected:
    y3g_integral
    describe_kind(int kind, volume_t const& v, double const* grid, int ngrid) const
    {
      y3g_integral ig;
      if (y3g_make_integral_volume(&p_, &spec_, kind, v.lows.data(), v.highs.data(),
                                   grid, ngrid, &ig) != Y3G_OK)
        throw std::runtime_error(y3g_last_error());
      return ig;
    }
    y3g_params p_{};
    y3g_table_spec spec_{}, base_spec_{};
    std::vector<double> hmf_tab_, da_tab_;

  private:
    static void
    require_uniform_(std::vector<double> const& x, std::string const& what)
    {
      if (x.size() < 2) throw std::runtime_error(what + " needs at least two points");
      double const step = (x.back() - x.front()) / double(x.size() - 1);
      if (!(step > 0.0)) throw std::runtime_error(what + " must increase");
      for (std::size_t i = 1; i < x.size(); ++i)
        if (std::fabs((x[i] - x[i - 1]) - step) > 1e-6 * step)
          throw std::runtime_error(what + " is not uniformly spaced; the device "
                                          "tables need uniform grids");
    }

    void
    read_hmf_table_(cosmosis::DataBlock& sample)
    {
      std::string const C = "cosmological_parameters", F = "mass_function";
      auto const mh = get_vector_double(sample, F, "m_h");
      auto const z = get_vector_double(sample, F, "z");
      auto const nd = sample.view<cosmosis::ndarray<double>>(F, "dndlnmh");
      double const om = p_.omega_m;
      double const onu = optional_double(sample, C, "omega_nu", 0.0);
      std::vector<double> x(mh.size());
      for (std::size_t i = 0; i < mh.size(); ++i) x[i] = std::log(mh[i] * (om - onu));
      require_uniform_(x, "mass_function ln(m_h (omega_M - omega_nu))");
      require_uniform_(z, "mass_function/z");
      if (nd.ndims() != 2 || nd.extents()[0] != z.size() || nd.extents()[1] != mh.size())
        throw std::runtime_error("mass_function/dndlnmh must have shape (n_z, n_m)");
      if (mh.size() * z.size() > 6400)
        throw std::runtime_error("mass_function table larger than 6400 entries");
      spec_.lnm_lo = x.front();
      spec_.lnm_hi = x.back();
      spec_.n_lnm = int(x.size());
      spec_.z_lo = z.front();
      spec_.z_hi = z.back();
      spec_.n_z = int(z.size());
      hmf_tab_.assign(nd.begin(), nd.end());
    }

    void
    read_da_table_(cosmosis::DataBlock& sample)
    {
      auto const z = get_vector_double(sample, "distances", "z");
      auto const da = get_vector_double(sample, "distances", "d_a");
      require_uniform_(z, "distances/z");
      if (std::fabs(z.front()) > 1e-12 || da.size() != z.size() || z.size() > 512)
        throw std::runtime_error("distances/{z, d_a} must start at z = 0, match in "
                                 "length and hold at most 512 points");
      spec_.da_zmax = z.back();
      spec_.n_da = int(z.size());
      da_tab_ = da;
    }
  };

  // ------------------------------------------------------ evaluator modules

  // A module whose class does a fixed quadrature of its own per sample (on
  // the device, in set_sample) and is then queried per grid point; each of
  // its K outputs is one 1-D array "vals" in its own section (reference
  // CosmoSISScalarEvaluatorModule, src/utils/CosmoSISScalarEvaluatorModule.hh
  // :9-26, 111-136).  The class is constructed from the config block and the
  // device runtime of this process.
  template <class I>
  class HIPEvaluatorModule {
  public:
    explicit HIPEvaluatorModule(cosmosis::DataBlock& cfg)
      : rank_(get_rank_info()),
        device_(rank_.local_rank),
        integrand_(cfg, device_.get()),
        grid_(I::make_grid_points(cfg))
    {}

    void
    execute(cosmosis::DataBlock& sample)
    {
      integrand_.set_sample(sample);
      constexpr std::size_t K = I::n_outputs;
      std::size_t const n = grid_.size();
      std::array<std::vector<double>, K> buf;
      for (auto& b : buf) b.assign(n, 0.0);
      for (std::size_t i = 0; i != n; ++i) {
        auto const v = integrand_.evaluate(grid_.points[i]);
        for (std::size_t k = 0; k != K; ++k) buf[k][i] = v[k];
      }
      auto const sections = I::output_sections();
      for (std::size_t k = 0; k != K; ++k)
        if (sample.put_val(sections[k], "vals", cosmosis::ndarray<double>(buf[k], {n})) !=
            DBS_SUCCESS)
          throw std::runtime_error(std::string(sections[k]) +
                                   "/vals could not be put into the sample");
    }

  private:
    rank_info rank_;
    DeviceBinding device_;
    I integrand_;
    typename I::grid_t grid_;
  };

} // namespace y3g_cosmosis

// The three CosmoSIS entry points for module class `klass` (reference:
// src/utils/cuda_module_macros.cuh:20-45 for the integration modules,
// src/utils/module_macros.hh:76-99 for the evaluators), with exceptions
// stopped at the C boundary.
#define DEFINE_COSMOSIS_HIP_INTEGRATION_MODULE(klass)                          \
  using y3g_module_type = y3g_cosmosis::HIPIntegrationModule<klass>;           \
  Y3G_COSMOSIS_ENTRY_POINTS(klass)

#define DEFINE_COSMOSIS_HIP_EVALUATOR_MODULE(klass)                            \
  using y3g_module_type = y3g_cosmosis::HIPEvaluatorModule<klass>;             \
  Y3G_COSMOSIS_ENTRY_POINTS(klass)

#define Y3G_COSMOSIS_ENTRY_POINTS(klass)                                       \
  extern "C" {                                                                 \
  void*                                                                        \
  setup(cosmosis::DataBlock* cfg)                                              \
  {                                                                            \
    if (!cfg) return nullptr;                                                  \
    try {                                                                      \
      return new y3g_module_type(*cfg);                                        \
    }                                                                          \
    catch (std::exception const& e) {                                          \
      std::cerr << "\nDuring construction of the module " << klass::module_label() \
                << " an exception was thrown:\n" << e.what() << '\n';          \
    }                                                                          \
    catch (...) {                                                              \
      std::cerr << "\nUnknown exception while constructing the module "        \
                << klass::module_label() << "\n";                              \
    }                                                                          \
    return nullptr;                                                            \
  }                                                                            \
                                                                               \
  DATABLOCK_STATUS                                                             \
  execute(cosmosis::DataBlock* sample, void* module)                           \
  {                                                                            \
    if (!sample) return DBS_DATABLOCK_NULL;                                    \
    if (!module) return DBS_LOGIC_ERROR;                                       \
    try {                                                                      \
      static_cast<y3g_module_type*>(module)->execute(*sample);                 \
      return DBS_SUCCESS;                                                      \
    }                                                                          \
    catch (std::exception const& e) {                                          \
      std::cerr << "\nModule " << klass::module_label() << " execute: "        \
                << e.what() << '\n';                                           \
    }                                                                          \
    catch (...) {                                                              \
      std::cerr << "\nModule " << klass::module_label()                        \
                << " execute: unknown exception\n";                            \
    }                                                                          \
    return DBS_LOGIC_ERROR;                                                    \
  }                                                                            \
                                                                               \
  int                                                                          \
  cleanup(void* module)                                                        \
  {                                                                            \
    delete static_cast<y3g_module_type*>(module);                              \
    return 0;                                                                  \
  }                                                                            \
  }

#endif
