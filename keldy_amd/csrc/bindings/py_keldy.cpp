// py_keldy.cpp -- Python bindings of the host API with the reference's Python
// names (python/keldy/{common,warpers,impurity_oneband}_desc.py).
#include <pybind11/complex.h>
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "../host/integrator.hpp"
#include "../host/model.hpp"
#include "../host/qrng.hpp"
#include "../host/warpers.hpp"

namespace py = pybind11;
using namespace keldy;
using namespace keldy::impurity_oneband;
using namespace keldy::warpers;

namespace {

// model_param_t from a dict with the reference's keys and defaults
// (impurity_oneband_desc.py:576-631).
model_param_t params_from_dict(py::dict d) {
  model_param_t p;
  static const char *keys[] = {"beta",          "bias_V",         "eps_d",    "Gamma",           "alpha",
                               "half_bandwidth", "bath_type",     "time_max", "nr_time_points_gf", "ft_method"};
  for (auto item : d) {
    const std::string k = py::str(item.first);
    bool known = false;
    for (auto kk : keys) known = known || k == kk;
    if (!known) throw keldy_error("model_param_t: unknown key '" + k + "'");
  }
  auto get = [&](const char *k, auto &v) {
    if (d.contains(k)) v = d[k].cast<std::decay_t<decltype(v)>>();
  };
  get("beta", p.beta);
  get("bias_V", p.bias_V);
  get("eps_d", p.eps_d);
  get("Gamma", p.Gamma);
  get("alpha", p.alpha);
  get("half_bandwidth", p.half_bandwidth);
  get("bath_type", p.bath_type);
  get("time_max", p.time_max);
  get("nr_time_points_gf", p.nr_time_points_gf);
  get("ft_method", p.ft_method);
  return p;
}

py::dict params_to_dict(model_param_t const &p) {
  py::dict d;
  d["beta"] = p.beta;
  d["bias_V"] = p.bias_V;
  d["eps_d"] = p.eps_d;
  d["Gamma"] = p.Gamma;
  d["alpha"] = p.alpha;
  d["half_bandwidth"] = p.half_bandwidth;
  d["bath_type"] = p.bath_type;
  d["time_max"] = p.time_max;
  d["nr_time_points_gf"] = p.nr_time_points_gf;
  d["ft_method"] = p.ft_method;
  return d;
}

py::array_t<dcomplex> table_array(g0_model const &m, std::vector<dcomplex> const &t) {
  py::array_t<dcomplex> a({m.n_orb, m.n_orb, m.n_pts});
  std::copy(t.begin(), t.end(), a.mutable_data());
  return a;
}

template <typename W, typename C> void bind_map_methods(C &c) {
  c.def("map_reverse", &W::map_reverse, py::arg("li_vec"))
      .def("map_forward", &W::map_forward, py::arg("ui_vec"))
      .def("ui_from_li", &W::ui_from_li, py::arg("li_vec"))
      .def("li_from_ui", &W::li_from_ui, py::arg("ui_vec"))
      .def("jacobian_reverse", &W::jacobian_reverse, py::arg("li_vec"))
      .def("jacobian_forward", &W::jacobian_forward, py::arg("ui_vec"));
}

template <typename W> void bind_product_1d(py::module_ &m, const char *name) {
  py::class_<W> c(m, name);
  c.def(py::init<>())
      .def("reserve", &W::reserve)
      .def("size", &W::size)
      .def("emplace_back", &W::emplace_back, py::arg("w"))
      .def("__call__", &W::operator());
  bind_map_methods<W>(c);
}

template <typename T, typename C> void bind_integrator_methods(C &c);

template <typename T> void bind_integrator(py::module_ &m, const char *name) {
  py::class_<T, integrator_direct> c(m, name);
  c.def(py::init<g0_model, double, int, double>(), py::arg("model"), py::arg("time"), py::arg("order"),
        py::arg("cutoff_integrand"))
      .def(py::init([](py::dict params, double time, int order, double cutoff) {
             return new T(params_from_dict(params), time, order, cutoff);
           }),
           py::arg("params"), py::arg("time"), py::arg("order"), py::arg("cutoff_integrand"))
      .def("reduce_result", &T::reduce_result)
      .def("get_integrand", &T::get_integrand);
  bind_integrator_methods<T>(c);
}

template <typename T, typename C> void bind_integrator_methods(C &c) {
  c.def_readwrite("warper", &T::warper)
      .def(
          "evaluate_warped_integrand",
          [](T &s, std::vector<double> li, int start, bool keep) { return s.evaluate_warped_integrand(li, start, keep); },
          py::arg("li_vec"), py::arg("start_domain_nr"), py::arg("keep_u_hypercube") = true)
      .def(
          "evaluate_warped_integrand", [](T &s, std::vector<double> li) { return s.evaluate_warped_integrand(li); },
          py::arg("li_vec"))
      .def(
          "evaluate_warped_integrand_batch",
          [](T &s, py::array_t<double, py::array::c_style | py::array::forcecast> l, bool keep) {
            if (l.ndim() != 2 || l.shape(1) != s.get_dimension())
              throw keldy_error("expected an array of shape (n, order)");
            const long n = long(l.shape(0));
            py::array_t<dcomplex> out(n);
            py::array_t<int> dom(n);
            py::array_t<double> jac(n);
            {
              py::gil_scoped_release rel;
              s.evaluate_warped_integrand_batch(l.data(), n, out.mutable_data(), dom.mutable_data(),
                                                jac.mutable_data(), keep);
            }
            return py::make_tuple(out, dom, jac);
          },
          py::arg("li"), py::arg("keep_u_hypercube") = true)
      .def("run", &T::run, py::arg("nr_steps"), py::call_guard<py::gil_scoped_release>())
      .def("reset_rng", &T::reset_rng, py::arg("rng_name"), py::arg("rng_state_seed"), py::arg("do_shift") = false,
           py::arg("do_scramble") = false, py::arg("rng_seed_shift") = 0)
      .def("reduce_nr_points_run", &T::reduce_nr_points_run)
      .def("reduce_nr_points_in_domain", &T::reduce_nr_points_in_domain)
      .def("reduce_nr_fallback", &T::reduce_nr_fallback)
      .def("get_warper", &T::get_warper)
      .def(
          "set_precision",
          [](T &s, std::string const &p) {
            if (p == "fp64")
              s.set_precision(KD_PREC_FP64);
            else if (p == "fp32")
              s.set_precision(KD_PREC_FP32_LU);
            else if (p == "fp32_all")
              s.set_precision(KD_PREC_FP32_LU_ALL);
            else
              throw keldy_error("precision must be 'fp64', 'fp32' (FP64 LU kept from order 10) or 'fp32_all'");
          },
          py::arg("precision"))
      .def(
          "set_algorithm",
          [](T &s, std::string const &a) {
            int v;
            if (a == "tree")
              v = KD_ALGO_TREE;
            else if (a == "config")
              v = KD_ALGO_CONFIG;
            else if (a == "pair")
              v = KD_ALGO_PAIR;
            else if (a == "split")
              v = KD_ALGO_SPLIT;
            else if (a == "wsplit")
              v = KD_ALGO_WSPLIT;
            else if (a == "ctree")
              v = KD_ALGO_CTREE;
            else if (a == "default")
              v = KD_ALGO_DEFAULT;
            else
              throw keldy_error("algorithm must be 'default', 'config', 'pair', 'split', 'wsplit', 'ctree' or 'tree'");
            s.ctx().check(kd_set_algorithm(s.ctx().get(), v));
          },
          py::arg("algorithm"), "Keldysh-sum algorithm: 'default' (ctree at orders 6..10, config otherwise), 'config', 'pair' "
          "(orders >= 7), 'split' or 'wsplit' (orders 6..10), 'ctree' (the time-ordered tree, orders 6..10) or 'tree' "
          "(orders >= 6)")
      .def("get_algorithm",
           [](T &s) {
             int32_t v = KD_ALGO_CONFIG;
             s.ctx().check(kd_get_algorithm(s.ctx().get(), &v));
             return std::string(v == KD_ALGO_TREE    ? "tree"
                                : v == KD_ALGO_PAIR  ? "pair"
                                : v == KD_ALGO_SPLIT ? "split"
                                : v == KD_ALGO_WSPLIT ? "wsplit"
                                : v == KD_ALGO_CTREE ? "ctree"
                                                     : "config");
           })
      .def("set_timing", &T::set_timing, py::arg("level"),
           "what last_timing measures: 0 nothing, 1 the whole call, 2 also the kernels (default); each "
           "timing event costs wall time on small runs")
      .def("last_timing", &T::last_timing)
      .def("flops_per_point", &T::flops_per_point)
      .def("get_integrand_dimension", &T::get_dimension)
      .def("make_warper_from_proj_device", &T::make_warper_from_proj_device, py::arg("num_bins"),
           py::arg("nr_samples"), py::arg("sigma"), py::arg("keep_u_hypercube") = false,
           py::call_guard<py::gil_scoped_release>(),
           "make_warper_from_proj with the pilot evaluations and the per-axis histograms on the device")
      .def("autotune", &T::autotune, py::arg("sample_points") = 1 << 18, py::call_guard<py::gil_scoped_release>(),
           "tune the Keldysh-sum kernel's launch shape: [lanes per point, waves per CU, LU ms per 2^20 points]")
      .def(
          "train_flow_warper",
          [](T &s, int n_layers, long nr_samples, int steps, int batch, double lr, double lr_final, int seed_shift,
             uint64_t init_seed, double weight_power) {
            std::pair<warper_flow_t, std::vector<double>> r;
            {
              py::gil_scoped_release rel;
              r = s.train_flow_warper(n_layers, nr_samples, steps, batch, lr, lr_final, seed_shift, init_seed,
                                      weight_power);
            }
            py::array_t<double> losses(py::ssize_t(r.second.size()));
            std::copy(r.second.begin(), r.second.end(), losses.mutable_data());
            return py::make_tuple(r.first, losses);
          },
          py::arg("n_layers") = 6, py::arg("nr_samples") = 1 << 20, py::arg("steps") = 2000,
          py::arg("batch") = 1 << 14, py::arg("lr") = 1e-3, py::arg("lr_final") = 1e-5,
          py::arg("rng_seed_shift") = 1000003, py::arg("init_seed") = 0, py::arg("weight_power") = 1.0,
          "Train a learned coupling-flow warper on |warped integrand| samples of the current chain "
          "(returns (WarperFlowT, per-step loss)); append it last to the warper train.");
}

} // namespace

PYBIND11_MODULE(_keldy_host, mod) {
  mod.doc() = "keldy_amd host API (C++), device path through libkeldy_amd.so";
  py::register_exception<keldy_error>(mod, "KeldyError", PyExc_RuntimeError);

  // ---------------------------------------------------------------- common
  py::module_ common = mod.def_submodule("common");
  py::class_<contour_pt_t>(common, "ContourPtT")
      .def(py::init<>())
      .def_readwrite("time", &contour_pt_t::time)
      .def_property(
          "k_idx", [](contour_pt_t const &c) { return int(c.k_idx); },
          [](contour_pt_t &c, int k) { c.k_idx = keldysh_idx_t(k); })
      .def_readwrite("timesplit_n", &contour_pt_t::timesplit_n);
  common.def("compare_3way", &compare_3way, py::arg("a"), py::arg("b"));
  py::class_<sobol>(common, "Sobol")
      .def(py::init<int, int, int>(), py::arg("dim"), py::arg("rng_state_seed"), py::arg("log_max_points_") = 31)
      .def(py::init<int, int, bool, bool, int, int>(), py::arg("dim"), py::arg("rng_state_seed"), py::arg("do_shift"),
           py::arg("do_scramble"), py::arg("rng_seed_shift"), py::arg("log_max_points_") = 31)
      .def("__call__", &sobol::operator())
      .def("seed", &sobol::seed, py::arg("k"))
      .def("discard", &sobol::discard, py::arg("nr_discard"))
      .def_property_readonly("state", &sobol::state)
      .def("generating_tables", [](sobol const &s) { return py::make_tuple(s.csr(), s.shfr()); });
  common.def("jk2008_matrices", &jk2008_matrices, py::arg("dim"), py::arg("m") = 31);

  // --------------------------------------------------------------- warpers
  py::module_ wm = mod.def_submodule("warpers");
  {
    py::class_<warper_identity_t> c(wm, "WarperIdentityT");
    c.def(py::init<>());
    bind_map_methods<warper_identity_t>(c);
  }
  {
    py::class_<warper_plasma_uv_t> c(wm, "WarperPlasmaUvT");
    c.def(py::init<double>(), py::arg("t_max_"));
    bind_map_methods<warper_plasma_uv_t>(c);
  }
  {
    py::class_<warper_product_1d_simple_t> c(wm, "WarperProduct1dSimpleT");
    c.def(py::init<>())
        .def(py::init<std::function<double(double)>, std::function<double(double)>, std::function<double(double)>,
                      double, bool>(),
             py::arg("f1_"), py::arg("f1_integrated_"), py::arg("f1_integrated_inverse_"), py::arg("domain_u_max_"),
             py::arg("do_domain_checks") = true)
        .def("__call__", &warper_product_1d_simple_t::operator());
    bind_map_methods<warper_product_1d_simple_t>(c);
  }
  {
    py::class_<warper_product_1d_simple_interp_nearest_t> c(wm, "WarperProduct1dSimpleInterpNearestT");
    c.def(py::init<>())
        .def(py::init<std::function<double(double)> const &, double, int>(), py::arg("f1_"), py::arg("domain_u_max_"),
             py::arg("nr_sample_points_"))
        .def(py::init<std::vector<double>, std::vector<double>>(), py::arg("times_u_pts_"), py::arg("f1_pts_"))
        .def("__call__", &warper_product_1d_simple_interp_nearest_t::operator())
        .def("knots",
             [](warper_product_1d_simple_interp_nearest_t const &w) {
               return py::make_tuple(w.m.knot_times(), w.m.knot_f1(), w.m.knot_F(), w.m.cut_index());
             },
             "(times, normalised f1, normalised negated CDF, cut index) of the tabulated map");
    bind_map_methods<warper_product_1d_simple_interp_nearest_t>(c);
  }
  {
    py::class_<warper_product_1d_simple_interp_hybrid_t> c(wm, "WarperProduct1dSimpleInterpHybridT");
    c.def(py::init<>())
        .def(py::init<std::function<double(double)>, double, int>(), py::arg("f1_"), py::arg("domain_u_max_"),
             py::arg("nr_sample_points_"))
        .def("__call__", &warper_product_1d_simple_interp_hybrid_t::operator());
    bind_map_methods<warper_product_1d_simple_interp_hybrid_t>(c);
  }
  bind_product_1d<warper_product_1d_t>(wm, "WarperProduct1dT");
  bind_product_1d<warper_product_1d_interp_hybrid_t>(wm, "WarperProduct1dInterpHybridT");
  {
    py::class_<warper_product_1d_interp_nearest_t> c(wm, "WarperProduct1dInterpNearestT");
    c.def(py::init<>())
        .def("reserve", &warper_product_1d_interp_nearest_t::reserve)
        .def("size", &warper_product_1d_interp_nearest_t::size)
        .def("emplace_back", &warper_product_1d_interp_nearest_t::emplace_back, py::arg("w"))
        .def("get_1d_warper", &warper_product_1d_interp_nearest_t::get_1d_warper, py::arg("axis"))
        .def("__call__", &warper_product_1d_interp_nearest_t::operator());
    bind_map_methods<warper_product_1d_interp_nearest_t>(c);
  }
  {
    py::class_<warper_flow_t> c(wm, "WarperFlowT");
    c.def(py::init([](int order, int n_layers, py::array_t<float, py::array::c_style | py::array::forcecast> p,
                      int n_cond, double cond) {
            return warper_flow_t(order, n_layers, std::vector<float>(p.data(), p.data() + p.size()), n_cond, cond);
          }),
          py::arg("order"), py::arg("n_layers"), py::arg("params"), py::arg("n_cond") = 0, py::arg("cond") = 0.0)
        .def("order", &warper_flow_t::order)
        .def("n_layers", &warper_flow_t::n_layers)
        .def("n_cond", &warper_flow_t::n_cond)
        .def("cond", &warper_flow_t::cond)
        .def("with_cond", &warper_flow_t::with_cond, py::arg("cond"),
             "the same (conditioned) flow at another conditioning value")
        .def("params", [](warper_flow_t const &w) {
          py::array_t<float> a(py::ssize_t(w.params().size()));
          std::copy(w.params().begin(), w.params().end(), a.mutable_data());
          return a;
        });
    bind_map_methods<warper_flow_t>(c);
  }
  {
    py::class_<warper_projection_t> c(wm, "WarperProjectionT");
    c.def(py::init([](integrator_direct &computer, int order, int num_bins, int nr_samples, double sigma,
                      bool optimize_sigma) {
            if (order != computer.get_dimension()) throw keldy_error("order differs from the integrator's dimension");
            return computer.make_warper_projection_device(num_bins, nr_samples, sigma, optimize_sigma, false);
          }),
          py::arg("warped_integrand"), py::arg("order"), py::arg("num_bins"), py::arg("nr_samples"), py::arg("sigma"),
          py::arg("optimize_sigma") = true,
          "The pilot on the device: the integrator's warped integrand (current chain, without the positive-time "
          "restriction) at the first nr_samples unshifted Sobol points.")
        .def(py::init<std::function<dcomplex(vec)>, int, int, int, double, bool>(), py::arg("warped_integrand"),
             py::arg("order"), py::arg("num_bins"), py::arg("nr_samples"), py::arg("sigma"),
             py::arg("optimize_sigma") = true)
        .def_static(
            "from_samples",
            [](py::array_t<double, py::array::c_style | py::array::forcecast> l, std::vector<double> const &vals,
               int order, int num_bins, double sigma, bool optimize_sigma) {
              if (l.ndim() != 2 || l.shape(1) != order) throw keldy_error("l must be [n][order]");
              std::vector<vec> pts(size_t(l.shape(0)));
              for (py::ssize_t i = 0; i < l.shape(0); i++) pts[size_t(i)].assign(l.data(i, 0), l.data(i, 0) + order);
              return warper_projection_t::from_samples(pts, vals, order, num_bins, sigma, optimize_sigma);
            },
            py::arg("l"), py::arg("abs_values"), py::arg("order"), py::arg("num_bins"), py::arg("sigma"),
            py::arg("optimize_sigma") = true)
        .def("get_xi", &warper_projection_t::get_xi, py::arg("axis"))
        .def("get_xi_counts", &warper_projection_t::get_xi_counts, py::arg("axis"))
        .def("get_xi_bin_coord", &warper_projection_t::get_xi_bin_coord, py::arg("axis"))
        .def("get_sigmas", &warper_projection_t::get_sigmas)
        .def("get_fi", &warper_projection_t::get_fi, py::arg("axis"))
        .def("get_fi_integrated", &warper_projection_t::get_fi_integrated, py::arg("axis"))
        .def("get_fi_integrated_inverse", &warper_projection_t::get_fi_integrated_inverse, py::arg("axis"))
        .def("order", &warper_projection_t::order)
        .def("num_bins", &warper_projection_t::num_bins);
    bind_map_methods<warper_projection_t>(c);
  }
  {
    py::class_<warper_train_t> c(wm, "WarperTrainT");
    c.def(py::init<>())
        .def("emplace_back", &warper_train_t::emplace_back<warper_flow_t>, py::arg("w"))
        .def("emplace_back", &warper_train_t::emplace_back<warper_identity_t>, py::arg("w"))
        .def("emplace_back", &warper_train_t::emplace_back<warper_plasma_uv_t>, py::arg("w"))
        .def("emplace_back", &warper_train_t::emplace_back<warper_product_1d_simple_t>, py::arg("w"))
        .def("emplace_back", &warper_train_t::emplace_back<warper_product_1d_simple_interp_nearest_t>, py::arg("w"))
        .def("emplace_back", &warper_train_t::emplace_back<warper_product_1d_simple_interp_hybrid_t>, py::arg("w"))
        .def("emplace_back", &warper_train_t::emplace_back<warper_product_1d_t>, py::arg("w"))
        .def("emplace_back", &warper_train_t::emplace_back<warper_product_1d_interp_nearest_t>, py::arg("w"))
        .def("emplace_back", &warper_train_t::emplace_back<warper_product_1d_interp_hybrid_t>, py::arg("w"))
        .def("emplace_back", &warper_train_t::emplace_back<warper_projection_t>, py::arg("w"))
        .def("reserve", &warper_train_t::reserve)
        .def("clear", &warper_train_t::clear)
        .def("size", &warper_train_t::size)
        .def("map_reverse", py::overload_cast<vec const &, int, int>(&warper_train_t::map_reverse, py::const_),
             py::arg("li_vec"), py::arg("start_domain_nr"), py::arg("end_domain_nr"))
        .def("map_forward", py::overload_cast<vec const &, int, int>(&warper_train_t::map_forward, py::const_),
             py::arg("ui_vec"), py::arg("start_domain_nr"), py::arg("end_domain_nr"))
        .def("map_reverse", py::overload_cast<vec const &>(&warper_train_t::map_reverse, py::const_),
             py::arg("li_vec"))
        .def("map_forward", py::overload_cast<vec const &>(&warper_train_t::map_forward, py::const_),
             py::arg("ui_vec"))
        .def("jacobian_reverse", &warper_train_t::jacobian_reverse, py::arg("li_vec"))
        .def("jacobian_forward", &warper_train_t::jacobian_forward, py::arg("ui_vec"))
        .def("ui_from_li", &warper_train_t::ui_from_li, py::arg("li_vec"))
        .def("li_from_ui", &warper_train_t::li_from_ui, py::arg("ui_vec"));
  }
  wm.def("vi_from_ui", &vi_from_ui, py::arg("t_max"), py::arg("u_times"));
  wm.def("ui_from_vi", &ui_from_vi, py::arg("t_max"), py::arg("v_times"));
  wm.def("make_product_1d_simple_exponential", &make_product_1d_simple_exponential, py::arg("domain_u_max"),
         py::arg("w_scale"));
  wm.def("make_product_1d_simple_inverse", &make_product_1d_simple_inverse, py::arg("domain_u_max"),
         py::arg("w_scale"));
  wm.def("make_product_1d_simple_inverse_square", &make_product_1d_simple_inverse_square, py::arg("domain_u_max"),
         py::arg("w_scale"));
  wm.def("make_product_1d_simple_inverse_power", &make_product_1d_simple_inverse_power, py::arg("power"),
         py::arg("domain_u_max"), py::arg("w_scale"));
  wm.def("make_product_1d_simple_uniform", &make_product_1d_simple_uniform, py::arg("domain_u_max"));
  wm.def("make_product_1d_inverse_cube_alternate", &make_product_1d_inverse_cube_alternate, py::arg("order"),
         py::arg("time"), py::arg("warper_scale"));
  wm.def("make_product_1d_inverse_cube_alternate_interp_nearest",
         &make_product_1d_inverse_cube_alternate_interp_nearest, py::arg("order"), py::arg("time"),
         py::arg("warper_scale"), py::arg("nr_sample_points_warper"));
  wm.def("make_product_1d_inverse_cube_alternate_interp_hybrid", &make_product_1d_inverse_cube_alternate_interp_hybrid,
         py::arg("order"), py::arg("time"), py::arg("warper_scale"), py::arg("nr_sample_points_warper"));
  wm.def("make_warper_from_proj", &make_warper_from_proj, py::arg("warped_integrand"), py::arg("order"),
         py::arg("num_bins"), py::arg("nr_samples"), py::arg("sigma"));
  wm.def(
      "make_warper_from_proj_samples",
      [](py::array_t<double, py::array::c_style | py::array::forcecast> l,
         py::array_t<double, py::array::c_style | py::array::forcecast> vals, int order, int num_bins, double sigma) {
        if (l.ndim() != 2 || l.shape(1) != order || vals.ndim() != 1 || vals.shape(0) != l.shape(0))
          throw keldy_error("expected l of shape (n, order) and values of shape (n,)");
        std::vector<vec> pts(size_t(l.shape(0)), vec(size_t(order)));
        for (ssize_t i = 0; i < l.shape(0); i++)
          for (int a = 0; a < order; a++) pts[size_t(i)][size_t(a)] = l.at(i, a);
        vec v(vals.data(), vals.data() + vals.shape(0));
        return make_warper_from_proj_samples(pts, v, order, num_bins, sigma);
      },
      py::arg("l"), py::arg("abs_values"), py::arg("order"), py::arg("num_bins"), py::arg("sigma"));
  wm.def("kernel_smoothing", &proj_details::kernel_smoothing, py::arg("y_in"), py::arg("sigma"));

  // ------------------------------------------------------ impurity_oneband
  py::module_ im = mod.def_submodule("impurity_oneband");
  py::class_<gf_index_t>(im, "GfIndexT")
      .def(py::init<>())
      .def(py::init<double, int, int>(), py::arg("time_"), py::arg("spin_"), py::arg("k_idx_"))
      .def(py::init<double, int, int, int>(), py::arg("time_"), py::arg("spin_"), py::arg("k_idx_"),
           py::arg("timesplit_n_"))
      .def(py::init<double, int, int, int, int>(), py::arg("time_"), py::arg("spin_"), py::arg("k_idx_"),
           py::arg("timesplit_n_"), py::arg("orbital_"))
      .def_readwrite("contour", &gf_index_t::contour)
      .def_property(
          "spin", [](gf_index_t const &g) { return int(g.spin); }, [](gf_index_t &g, int s) { g.spin = spin_t(s); })
      .def_readwrite("orbital", &gf_index_t::orbital);
  py::class_<g0_model_omega>(im, "G0ModelOmega")
      .def(py::init<>())
      .def(py::init([](py::dict p) { return g0_model_omega(params_from_dict(p)); }), py::arg("parameters"))
      .def("get_param", [](g0_model_omega const &m) { return params_to_dict(m.get_param()); })
      .def("get_param_alpha", &g0_model_omega::get_param_alpha)
      .def("mu_left", &g0_model_omega::mu_left)
      .def("mu_right", &g0_model_omega::mu_right)
      .def("g0_dot_R", &g0_model_omega::g0_dot_R, py::arg("omega"))
      .def("g0_dot_A", &g0_model_omega::g0_dot_A, py::arg("omega"))
      .def(
          "g0_dot_keldysh",
          [](g0_model_omega const &m, int a, int b, dcomplex w) {
            return m.g0_dot_keldysh(keldysh_idx_t(a), keldysh_idx_t(b), w);
          },
          py::arg("a"), py::arg("b"), py::arg("omega"))
      .def("g0_dot_lesser", &g0_model_omega::g0_dot_lesser, py::arg("omega"))
      .def("g0_dot_greater", &g0_model_omega::g0_dot_greater, py::arg("omega"))
      .def("bath_hybrid_R_left", &g0_model_omega::bath_hybrid_R_left, py::arg("omega"))
      .def("bath_hybrid_A_left", &g0_model_omega::bath_hybrid_A_left, py::arg("omega"))
      .def("bath_hybrid_K_left", &g0_model_omega::bath_hybrid_K_left, py::arg("omega"))
      .def("bath_hybrid_R_right", &g0_model_omega::bath_hybrid_R_right, py::arg("omega"))
      .def("bath_hybrid_A_right", &g0_model_omega::bath_hybrid_A_right, py::arg("omega"))
      .def("bath_hybrid_K_right", &g0_model_omega::bath_hybrid_K_right, py::arg("omega"))
      .def("g0_rightlead_dot_lesser", &g0_model_omega::g0_rightlead_dot_lesser, py::arg("omega"))
      .def("g0_rightlead_dot_greater", &g0_model_omega::g0_rightlead_dot_greater, py::arg("omega"));
  py::class_<g0_model>(im, "G0Model")
      .def(py::init<>())
      .def(py::init<g0_model_omega, bool>(), py::arg("model_omega_"), py::arg("make_dot_lead_"),
           py::call_guard<py::gil_scoped_release>())
      .def_static(
          "from_tables",
          [](g0_model_omega mo, bool dot_lead, double time_max, py::array_t<dcomplex, py::array::c_style> les,
             py::array_t<dcomplex, py::array::c_style> gre) {
            const int n_orb = les.ndim() == 3 ? int(les.shape(0)) : 1;
            std::vector<dcomplex> l(les.data(), les.data() + les.size()), g(gre.data(), gre.data() + gre.size());
            return g0_model(mo, dot_lead, n_orb, time_max, std::move(l), std::move(g));
          },
          py::arg("model_omega"), py::arg("make_dot_lead"), py::arg("time_max"), py::arg("lesser"),
          py::arg("greater"))
      .def_readwrite("model_omega", &g0_model::model_omega)
      .def_readonly("make_dot_lead", &g0_model::make_dot_lead)
      .def_property_readonly("g0_lesser", [](g0_model const &m) { return table_array(m, m.lesser); })
      .def_property_readonly("g0_greater", [](g0_model const &m) { return table_array(m, m.greater); })
      .def_property_readonly("lesser_ft_error", [](g0_model const &m) { return table_array(m, m.lesser_ft_error); })
      .def_property_readonly("greater_ft_error", [](g0_model const &m) { return table_array(m, m.greater_ft_error); })
      .def_property_readonly("time_mesh",
                             [](g0_model const &m) {
                               py::array_t<double> t(m.n_pts);
                               for (int k = 0; k < m.n_pts; k++) t.mutable_at(k) = m.time(k);
                               return t;
                             })
      .def_readonly("time_max", &g0_model::time_max)
      .def("lesser_at", [](g0_model const &m, double t, int a, int b) { return m.eval(m.lesser, a, b, t); },
           py::arg("t"), py::arg("a") = 0, py::arg("b") = 0)
      .def("greater_at", [](g0_model const &m, double t, int a, int b) { return m.eval(m.greater, a, b, t); },
           py::arg("t"), py::arg("a") = 0, py::arg("b") = 0)
      .def_static("hdf5_format", &g0_model::hdf5_format);
  py::class_<g0_keldysh_contour_t>(im, "G0KeldyshContourT")
      .def(py::init<g0_model>(), py::arg("model_"))
      .def_readwrite("model", &g0_keldysh_contour_t::model)
      .def("__call__", &g0_keldysh_contour_t::operator(), py::arg("a"), py::arg("b"), py::arg("internal_point") = true)
      .def("get_time_max", &g0_keldysh_contour_t::get_time_max);
  py::class_<integrand_g_direct>(im, "IntegrandGDirect")
      .def(py::init<g0_keldysh_contour_t, gf_index_t, gf_index_t, double>(), py::arg("g0_"), py::arg("external_A_"),
           py::arg("external_B_"), py::arg("cutoff_") = 0.)
      .def("__call__", &integrand_g_direct::operator(), py::arg("times"), py::arg("keep_u_hypercube") = true)
      .def("eval_no_det", &integrand_g_direct::eval_no_det, py::arg("times"), py::arg("keep_u_hypercube") = true)
      .def(
          "eval_batch",
          [](integrand_g_direct const &f, py::array_t<double, py::array::c_style | py::array::forcecast> u,
             bool keep) {
            if (u.ndim() != 2) throw keldy_error("expected an array of shape (n, order)");
            const long n = long(u.shape(0));
            py::array_t<dcomplex> out(n);
            py::array_t<int> dom(n);
            {
              py::gil_scoped_release rel;
              f.eval_batch(u.data(), n, int(u.shape(1)), keep, out.mutable_data(), dom.mutable_data());
            }
            return py::make_tuple(out, dom);
          },
          py::arg("u"), py::arg("keep_u_hypercube") = true);
  py::class_<integrator_direct>(im, "IntegratorDirect", "common base of the compute_* integrators");
  bind_integrator<compute_charge_Q_direct>(im, "ComputeChargeQDirect");
  bind_integrator<compute_current_J_direct>(im, "ComputeCurrentJDirect");
  // time-binned charge (compute_obs.hpp:85-101, wick_direct_time.hpp)
  py::class_<binner::binner_1d_t>(im, "Binner1", "binner_t<1> (binner.hpp:147-352)")
      .def("get_data", [](binner::binner_1d_t const &b) {
        py::array_t<dcomplex> a(py::ssize_t(b.data.size()));
        std::copy(b.data.begin(), b.data.end(), a.mutable_data());
        return a;
      })
      .def("get_nr_values_added", [](binner::binner_1d_t const &b) {
        py::array_t<long> a(py::ssize_t(b.nr_values_added.size()));
        std::copy(b.nr_values_added.begin(), b.nr_values_added.end(), a.mutable_data());
        return a;
      })
      .def("get_nr_values_dropped", [](binner::binner_1d_t const &b) { return b.nr_values_dropped; })
      .def("get_nr_bins", [](binner::binner_1d_t const &b, int axis) {
        if (axis != 0) throw keldy_error("Axis " + std::to_string(axis) + " is larger than the number of dimensions.");
        return b.nr_bins;
      }, py::arg("axis") = 0)
      .def("get_bin_coord", [](binner::binner_1d_t const &b, int axis) {
        if (axis != 0) throw keldy_error("Axis " + std::to_string(axis) + " is larger than the number of dimensions.");
        return b.get_bin_coord();
      }, py::arg("axis") = 0)
      .def("get_bin_size", [](binner::binner_1d_t const &b, int axis) {
        if (axis != 0) throw keldy_error("Axis " + std::to_string(axis) + " is larger than the number of dimensions.");
        return b.get_bin_size();
      }, py::arg("axis") = 0);
  py::class_<binner::sparse_binner_1d_t>(im, "SparseBinner1", "sparse_binner_t<1> (binner.hpp:59-140)")
      .def_readonly("data", &binner::sparse_binner_1d_t::data)
      .def("sum_moduli", &binner::sparse_binner_1d_t::sum_moduli);
  py::class_<integrand_g_direct_time>(im, "IntegrandGDirectTime")
      .def(py::init<g0_keldysh_contour_t, gf_index_t, gf_index_t, double>(), py::arg("g0_"), py::arg("external_A_"),
           py::arg("external_B_"), py::arg("cutoff_") = 0.)
      .def("__call__", &integrand_g_direct_time::operator(), py::arg("times"), py::arg("keep_u_hypercube") = true);
  {
    using T = compute_charge_Q_direct_time;
    py::class_<T, integrator_direct> c(im, "ComputeChargeQDirectTime");
    c.def(py::init<g0_model, double, int, int, double>(), py::arg("model"), py::arg("time"), py::arg("order"),
          py::arg("nr_time_slices"), py::arg("cutoff_integrand"))
        .def(py::init([](py::dict params, double time, int order, int slices, double cutoff) {
               return new T(params_from_dict(params), time, order, slices, cutoff);
             }),
             py::arg("params"), py::arg("time"), py::arg("order"), py::arg("nr_time_slices"),
             py::arg("cutoff_integrand"))
        .def("reduce_result", &T::reduce_result_binned)
        .def("get_integrand", &T::get_integrand_time)
        .def("evaluate_warped_integrand", &T::evaluate_warped_integrand_time, py::arg("li_vec"),
             py::arg("start_domain_nr"), py::arg("keep_u_hypercube") = true)
        .def(
            "evaluate_warped_integrand",
            [](T &s, std::vector<double> li) {
              return s.evaluate_warped_integrand_time(li, int(s.warper.size()), true);
            },
            py::arg("li_vec"));
    bind_integrator_methods<T>(c);
  }
  // kernel method (compute_obs.hpp:102-133, wick_kernel*.hpp)
  py::class_<binner::binner_1d_1_t>(im, "Binner_1_1", "binner_t<1,1> (binner.hpp:147-352): time x Keldysh index")
      .def("get_data", [](binner::binner_1d_1_t const &b) {
        py::array_t<dcomplex> a({py::ssize_t(b.nr_bins), py::ssize_t(b.nr_disc)});
        std::copy(b.data.begin(), b.data.end(), a.mutable_data());
        return a;
      })
      .def("get_nr_values_added", [](binner::binner_1d_1_t const &b) {
        py::array_t<long> a({py::ssize_t(b.nr_bins), py::ssize_t(b.nr_disc)});
        std::copy(b.nr_values_added.begin(), b.nr_values_added.end(), a.mutable_data());
        return a;
      })
      .def("get_nr_values_dropped", [](binner::binner_1d_1_t const &b) { return b.nr_values_dropped; })
      .def("get_nr_bins", [](binner::binner_1d_1_t const &b, int axis) {
        if (axis < 0 || axis > 1) throw keldy_error("Axis " + std::to_string(axis) + " is larger than the number of dimensions.");
        return axis == 0 ? b.nr_bins : b.nr_disc;
      }, py::arg("axis") = 0)
      .def("get_bin_coord", [](binner::binner_1d_1_t const &b, int axis) {
        if (axis == 1) throw keldy_error("Axis 1 is discreet.");
        if (axis != 0) throw keldy_error("Axis " + std::to_string(axis) + " is larger than the number of dimensions.");
        return b.get_bin_coord();
      }, py::arg("axis") = 0)
      .def("get_bin_size", [](binner::binner_1d_1_t const &b, int axis) {
        if (axis == 1) throw keldy_error("Axis 1 is discreet.");
        if (axis != 0) throw keldy_error("Axis " + std::to_string(axis) + " is larger than the number of dimensions.");
        return b.get_bin_size();
      }, py::arg("axis") = 0)
      .def("get_continuous_axes", [](binner::binner_1d_1_t const &b) {
        return std::vector<std::tuple<double, double, long, double>>{{b.xmin, b.xmax, b.nr_bins, b.get_bin_size()}};
      })
      .def("get_discreet_axes", [](binner::binner_1d_1_t const &b) { return std::vector<long>{b.nr_disc}; });
  py::class_<binner::sparse_binner_1d_1_t>(im, "SparseBinner_1_1", "sparse_binner_t<1,1> (binner.hpp:59-140)")
      .def_readonly("data", &binner::sparse_binner_1d_1_t::data)
      .def("sum_moduli", &binner::sparse_binner_1d_1_t::sum_moduli);
  py::class_<integrand_g_kernel>(im, "IntegrandGKernel")
      .def(py::init<g0_keldysh_contour_t, gf_index_t>(), py::arg("g0_"), py::arg("g_idx_X_"))
      .def("__call__", &integrand_g_kernel::operator(), py::arg("times"), py::arg("keep_u_hypercube") = true)
      .def(
          "eval_entries",
          [](integrand_g_kernel const &f, py::array_t<double, py::array::c_style | py::array::forcecast> u, bool keep) {
            if (u.ndim() != 2 || u.shape(1) < 1) throw keldy_error("expected an array of shape (n, order)");
            const long n = long(u.shape(0));
            const int order = int(u.shape(1));
            if (order > KD_MAX_ORDER) throw keldy_error("order exceeds the device maximum");
            py::array_t<dcomplex> out({py::ssize_t(n), py::ssize_t(1) << order, py::ssize_t(order)});
            py::array_t<int> dom(n);
            {
              py::gil_scoped_release rel;
              f.eval_entries(u.data(), n, order, keep, out.mutable_data(), dom.mutable_data());
            }
            return py::make_tuple(out, dom);
          },
          py::arg("u"), py::arg("keep_u_hypercube") = true,
          "entries [n][2^order][order]: configuration c, vertex i (time u_i, Keldysh index bit i of c)");
  py::class_<integrand_g_kernel_single_omega>(im, "IntegrandGKernelSingleOmega")
      .def(py::init<g0_keldysh_contour_t, gf_index_t, double>(), py::arg("g0_"), py::arg("g_idx_X_"), py::arg("omega_"))
      .def("__call__", &integrand_g_kernel_single_omega::operator(), py::arg("times"),
           py::arg("keep_u_hypercube") = true)
      .def(
          "eval_batch",
          [](integrand_g_kernel_single_omega const &f, py::array_t<double, py::array::c_style | py::array::forcecast> u,
             bool keep) {
            if (u.ndim() != 2) throw keldy_error("expected an array of shape (n, order)");
            const long n = long(u.shape(0));
            py::array_t<dcomplex> out(n);
            py::array_t<int> dom(n);
            {
              py::gil_scoped_release rel;
              f.eval_batch(u.data(), n, int(u.shape(1)), keep, out.mutable_data(), dom.mutable_data());
            }
            return py::make_tuple(out, dom);
          },
          py::arg("u"), py::arg("keep_u_hypercube") = true);
  {
    using T = compute_gf_kernel;
    py::class_<T, integrator_direct> c(im, "ComputeGfKernel");
    c.def(py::init<g0_model, double, int, int>(), py::arg("model"), py::arg("time"), py::arg("order"),
          py::arg("nr_bins") = 100)
        .def(py::init([](py::dict params, double time, int order, int nr_bins) {
               return new T(params_from_dict(params), time, order, nr_bins);
             }),
             py::arg("params"), py::arg("time"), py::arg("order"), py::arg("nr_bins") = 100)
        .def("reduce_result", &T::reduce_result_binned)
        .def("get_integrand", &T::get_integrand_kernel)
        .def("evaluate_warped_integrand", &T::evaluate_warped_integrand_kernel, py::arg("li_vec"),
             py::arg("start_domain_nr"), py::arg("keep_u_hypercube") = true)
        .def(
            "evaluate_warped_integrand",
            [](T &s, std::vector<double> li) {
              return s.evaluate_warped_integrand_kernel(li, int(s.warper.size()), true);
            },
            py::arg("li_vec"))
        .def(
            "evaluate_warped_entries_batch",
            [](T &s, py::array_t<double, py::array::c_style | py::array::forcecast> l, bool keep) {
              const int order = s.get_dimension();
              if (l.ndim() != 2 || l.shape(1) != order) throw keldy_error("expected an array of shape (n, order)");
              const long n = long(l.shape(0));
              py::array_t<dcomplex> out({py::ssize_t(n), py::ssize_t(1) << order, py::ssize_t(order)});
              py::array_t<int> dom(n);
              py::array_t<double> jac(n);
              {
                py::gil_scoped_release rel;
                s.evaluate_warped_entries_batch(l.data(), n, out.mutable_data(), dom.mutable_data(),
                                                jac.mutable_data(), keep);
              }
              return py::make_tuple(out, dom, jac);
            },
            py::arg("li"), py::arg("keep_u_hypercube") = true);
    bind_integrator_methods<T>(c);
  }
  {
    using T = compute_gf_kernel_single_omega;
    py::class_<T, integrator_direct> c(im, "ComputeGfKernelSingleOmega");
    c.def(py::init<g0_model, double, double, int>(), py::arg("model"), py::arg("time"), py::arg("omega"),
          py::arg("order"))
        .def(py::init([](py::dict params, double time, double omega, int order) {
               return new T(params_from_dict(params), time, omega, order);
             }),
             py::arg("params"), py::arg("time"), py::arg("omega"), py::arg("order"))
        .def("reduce_result", &T::reduce_result)
        .def("get_integrand", &T::get_integrand_omega);
    bind_integrator_methods<T>(c);
  }
  im.def(
      "run_batch",
      [](py::list integrators, std::vector<long> nr_steps) {
        std::vector<integrator_direct *> its;
        for (auto h : integrators) its.push_back(h.cast<integrator_direct *>());
        py::gil_scoped_release rel;
        return run_batch(its, nr_steps);
      },
      py::arg("integrators"), py::arg("nr_steps"),
      "run several integrators (e.g. orders 1..N) in one hipGraph launch; returns its wall time in ms");
  im.def(
      "train_flow_warper_conditioned",
      [](py::list computers, std::vector<double> conds, int n_layers, long samples_per_computer, int steps, int batch,
         double lr, double lr_final, int seed_shift, uint64_t init_seed, double weight_power) {
        std::vector<integrator_direct *> its;
        for (auto h : computers) its.push_back(h.cast<integrator_direct *>());
        std::pair<warper_flow_t, std::vector<double>> r;
        {
          py::gil_scoped_release rel;
          r = integrator_direct::train_flow_warper_conditioned(its, conds, n_layers, samples_per_computer, steps,
                                                               batch, lr, lr_final, seed_shift, init_seed,
                                                               weight_power);
        }
        py::array_t<double> losses(py::ssize_t(r.second.size()));
        std::copy(r.second.begin(), r.second.end(), losses.mutable_data());
        return py::make_tuple(r.first, losses);
      },
      py::arg("computers"), py::arg("conds"), py::arg("n_layers") = 6, py::arg("samples_per_computer") = 1 << 16,
      py::arg("steps") = 2000, py::arg("batch") = 1 << 12, py::arg("lr") = 1e-3, py::arg("lr_final") = 1e-5,
      py::arg("rng_seed_shift") = 1000003, py::arg("init_seed") = 0, py::arg("weight_power") = 1.0,
      "Train one time-conditioned coupling flow on the samples of several computers (conds[i] tags computer i's "
      "samples); returns (WarperFlowT at conds[0], per-step loss).  flow.with_cond(conds[i]) is computer i's warper.");
}
