// C++ check of the time-integrator headers (include/flecsolve/time-integrators)
// written the way the reference's own tests use its integrators
// (flecsolve/time-integrators/test/implicit.cc:27-135, explicit.cc:16-101):
// RK23 / RK45 on u' = lambda u, and the rate problem
// u' = lambda u with a direct stage solver, configured from the reference's
// implicit.cfg (bdf-2, bdf-5 sections), and the 2D heat equation of the
// reference's heat example (examples/heat_equation/implicit.cc) with its
// stage systems solved by a CG chosen through krylov_factory.
//
// Built twice: with -DBDF_HOST over host vectors (no device use; the CPU
// suite compares it with the oracle) and over device vectors (libfsx).
//
//   bdf_check rate <cfg> <lambda> <ic> <n>
//       runs the bdf-2 and bdf-5 sections; prints per section
//       "<section> steps S rejects R time T value V" (V = max of u, as %a)
//   bdf_check rateone <cfg> <section> <lambda> <ic> <n>
//       the rate problem with the file's only section; prints one such line
//   bdf_check explicit <cfg> <lambda> <ic> <n>
//       the reference's explicit test (time-integrators/test/explicit.cc):
//       RK23 and RK45 from the fixed and variable sections; prints
//       "<rk23|rk45>-<fixed|variable> steps S rejects R time T value V"
//   bdf_check heat <cfg> <n> <alpha> <out.bin>
//       heat on an n x n interior grid of (0,10)^2, u0 = 50 on [4,6]^2;
//       prints "heat steps S rejects R time T solves K iters I" and writes
//       the final solution (n*n doubles) to out.bin
// Exit status 0 = ran.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <tuple>
#include <vector>

#include "flecsolve/solvers/factory.hh"
#include "flecsolve/time-integrators/bdf.hh"
#include "flecsolve/time-integrators/operator_adapter.hh"
#include "flecsolve/time-integrators/rk45.hh"
#include "fsx_problems.h"

#ifdef BDF_HOST
#include "host_vector.hh"
#else
#include "flecsolve/matrices/device_csr.hh"
#include "flecsolve/vectors/device.hh"
#endif

using namespace flecsolve;
namespace ti = flecsolve::time_integrator;

#ifdef BDF_HOST
struct backend {
	using vec_t = vec::host<double>;
	vec_t make(std::size_t n) { return vec::make_host(n); }
	void upload(vec_t & v, const std::vector<double> & h) {
		std::copy(h.begin(), h.end(), v.data.p());
	}
	std::vector<double> download(const vec_t & v) {
		return std::vector<double>(v.data.p(), v.data.p() + v.data.size());
	}
	auto heat_op(std::vector<int32_t> rp, std::vector<int32_t> ci, std::vector<double> v) {
		return op::make_shared<ti::operator_adapter<mat::host_csr>>(std::move(rp), std::move(ci),
		                                                           std::move(v));
	}
};
#else
struct backend {
	using vec_t = vec::device<double>;
	backend() { fsx_check(fsx_ctx_create(0, &ctx), "ctx"); }
	vec_t make(std::size_t n) { return vec_t(ctx, n); }
	void upload(vec_t & v, const std::vector<double> & h) {
		fsx_check(fsx_vec_upload(ctx, v.handle(), h.data(), h.size()), "upload");
	}
	std::vector<double> download(const vec_t & v) {
		std::vector<double> h(v.data.size());
		fsx_check(fsx_vec_download(ctx, v.handle(), h.data(), h.size()), "download");
		return h;
	}
	auto heat_op(std::vector<int32_t> rp, std::vector<int32_t> ci, std::vector<double> v) {
		fsx_csr * F;
		fsx_check(fsx_csr_create(ctx, (int64_t)rp.size() - 1, (int64_t)rp.size() - 1,
		                         (int64_t)ci.size(), FSX_I32, rp.data(), ci.data(), FSX_F64,
		                         v.data(), &F),
		          "csr");
		return op::make_shared<ti::operator_adapter<mat::device_csr>>(
			mat::device_csr::adopt(ctx, F));
	}
	fsx_ctx * ctx = nullptr;
};
#endif

// time-integrators/test/implicit.cc:27-63
struct rate_parameters {
	double lambda;
	double gamma;
};

struct rate : op::base<rate_parameters> {
	explicit rate(double lambda) : op::base<rate_parameters>(rate_parameters{lambda, 1.}) {}
	template<class D, class R>
	void apply(const D & x, R & y) const {
		apply_rhs(x, y);
		y.axpy(-params.gamma, y, x);
	}
	template<class D, class R>
	void apply_rhs(const D & x, R & y) const { y.scale(params.lambda, x); }
	template<class V>
	bool is_valid(const V &) { return true; }
	double get_scaling() const { return params.gamma; }
	void set_scaling(double s) { params.gamma = s; }
	double get_rate() const { return params.lambda; }
};

// implicit.cc:67-85: x = min(b) / (1 - lambda gamma)
struct rate_solver : op::base<> {
	explicit rate_solver(op::handle<op::core<rate>> F) : F(F) {}
	template<class D, class R>
	solve_info apply(const D & b, R & x) const {
		const double rhs = b.min().get();
		const auto & op = F.get();
		x.set_scalar(rhs / (1. - op.get_rate() * op.get_scaling()));
		solve_info info;
		info.status = solve_info::stop_reason::converged_atol;
		return info;
	}
	op::handle<op::core<rate>> F;
};

// counts the stage solves and their iterations
template<class H>
struct counting : op::base<> {
	explicit counting(H h) : inner(std::move(h)) {}
	template<class B, class X>
	solve_info apply(const B & b, X & x) const {
		auto info = inner.get().apply(b, x);
		iters += info.iters;
		++solves;
		return info;
	}
	H inner;
	mutable long iters = 0;
	mutable long solves = 0;
};

// the time loop of implicit.cc:101-114 and examples/heat_equation/implicit.cc:41-55
template<class TI, class V>
void integrate(TI & t, V & x, V & xnew) {
	auto dt = t.get_current_dt();
	bool first_step = true;
	while (t.get_current_time() < t.get_final_time()) {
		t.advance(dt, first_step, x, xnew);
		const bool good = t.check_solution();
		if (good) {
			t.update();
			std::swap(x, xnew);
			first_step = false;
		}
		dt = t.get_next_dt(good);
	}
}

// implicit.cc:87-134: both sections read in one read_config call
static int run_rate(backend & be, const char * cfg, double lambda, double ic, std::size_t n) {
	auto F = op::make_shared<rate>(lambda);
	auto x = be.make(n), xnew = be.make(n);
	auto solver = op::make_shared<rate_solver>(F);
	auto [t2, t5] = std::apply(
		[&](auto &&... s) {
			return std::make_tuple(
				ti::bdf::integrator(ti::bdf::parameters(s, F, ti::bdf::make_work(x), solver))...);
		},
		read_config(cfg, ti::bdf::options("bdf-2"), ti::bdf::options("bdf-5")));
	auto run = [&](auto & t, const char * name) {
		x.set_scalar(ic);
		integrate(t, x, xnew);
		std::printf("%s steps %d rejects %d time %.17g value %a\n", name, t.get_current_step(),
		            t.num_step_rejects(), t.get_current_time(), (double)x.max().get());
	};
	run(t2, "bdf-2");
	run(t5, "bdf-5");
	return 0;
}

// one configuration file holding one section: that integrator on the rate
// problem (the other branches of the integrator, against the oracle)
static int run_rate_one(backend & be, const char * cfg, const char * section, double lambda,
                        double ic, std::size_t n) {
	auto F = op::make_shared<rate>(lambda);
	auto x = be.make(n), xnew = be.make(n);
	auto solver = op::make_shared<rate_solver>(F);
	auto s = read_config(cfg, ti::bdf::options(section));
	ti::bdf::integrator t(ti::bdf::parameters(s, F, ti::bdf::make_work(x), solver));
	x.set_scalar(ic);
	integrate(t, x, xnew);
	std::printf("%s steps %d rejects %d time %.17g value %a\n", section, t.get_current_step(),
	            t.num_step_rejects(), t.get_current_time(), (double)x.max().get());
	return 0;
}

// time-integrators/test/explicit.cc:16-23: f(x) = lambda x
struct explicit_rate_parameters {
	double lambda;
};
struct explicit_rate : op::base<explicit_rate_parameters> {
	explicit explicit_rate(double l) : op::base<explicit_rate_parameters>(explicit_rate_parameters{l}) {}
	template<class D, class R>
	void apply(const D & x, R & y) const { y.scale(params.lambda, x); }
};

// explicit.cc:45-101: the variable and fixed sections for RK23 and RK45
static int run_explicit(backend & be, const char * cfg, double lambda, double ic,
                        std::size_t n) {
	op::core<explicit_rate> F(lambda);
	auto x = be.make(n), xnew = be.make(n);
	auto [t23v, t23f] = std::apply(
		[&](auto &&... s) {
			return std::make_tuple(
				ti::rk23::integrator(ti::rk23::parameters(s, op::ref(F), ti::rk23::make_work(x)))...);
		},
		read_config(cfg, ti::rk23::options("variable"), ti::rk23::options("fixed")));
	auto [t45v, t45f] = std::apply(
		[&](auto &&... s) {
			return std::make_tuple(
				ti::rk45::integrator(ti::rk45::parameters(s, op::ref(F), ti::rk45::make_work(x)))...);
		},
		read_config(cfg, ti::rk45::options("variable"), ti::rk45::options("fixed")));
	auto run = [&](auto & t, const char * name) {
		x.set_scalar(ic);
		auto dt = t.get_current_dt();
		while (t.get_current_time() < t.get_final_time()) {
			t.advance(dt, x, xnew);
			const bool good = t.check_solution();
			if (good || t.fixed_dt()) {
				t.update();
				std::swap(x, xnew);
			}
			dt = t.get_next_dt(good);
		}
		std::printf("%s steps %d rejects %d time %.17g value %a\n", name, t.get_current_step(),
		            t.num_step_rejects(), t.get_current_time(), (double)x.max().get());
	};
	run(t23f, "rk23-fixed");
	run(t45f, "rk45-fixed");
	run(t23v, "rk23-variable");
	run(t45v, "rk45-variable");
	return 0;
}

static int run_heat(backend & be, const char * cfg, int n, double alpha, const char * out) {
	std::vector<int32_t> rp((std::size_t)n * n + 1), ci(fsx_gen_poisson2d_nnz(n));
	std::vector<double> v(ci.size());
	fsx_gen_poisson2d(n, rp.data(), ci.data(), v.data());
	const double h = 10.0 / (n + 1);
	const double c = alpha / (h * h);
	for (auto & e : v) e = -c * e; // f(u) = alpha Laplacian(u), Dirichlet 0
	std::vector<double> u0((std::size_t)n * n);
	for (int j = 0; j < n; ++j)
		for (int i = 0; i < n; ++i) {
			const double xx = (i + 1) * h, yy = (j + 1) * h;
			u0[(std::size_t)j * n + i] = (xx >= 4 && xx <= 6 && yy >= 4 && yy <= 6) ? 50.0 : 0.0;
		}
	auto [ts, ss] = read_config(cfg, ti::bdf::options("time-integrator"),
	                            krylov_factory::options("linear-solver"));
	auto F = be.heat_op(rp, ci, v);
	auto u = be.make(u0.size()), unew = be.make(u0.size());
	be.upload(u, u0);
	auto inner = krylov_factory::make_shared(ss, u, F);
	auto solver = op::make_shared<counting<decltype(inner)>>(inner);
	ti::bdf::integrator t(ti::bdf::parameters(ts, F, ti::bdf::make_work(u), solver));
	integrate(t, u, unew);
	const auto uh = be.download(u);
	std::ofstream(out, std::ios::binary)
		.write(reinterpret_cast<const char *>(uh.data()), (std::streamsize)(uh.size() * 8));
	std::printf("heat steps %d rejects %d time %.17g solves %ld iters %ld\n",
	            t.get_current_step(), t.num_step_rejects(), t.get_current_time(),
	            solver.get().solves, solver.get().iters);
	return 0;
}

int main(int argc, char ** argv) {
	try {
		backend be;
		if (argc == 6 && !std::strcmp(argv[1], "rate"))
			return run_rate(be, argv[2], std::atof(argv[3]), std::atof(argv[4]),
			                (std::size_t)std::atol(argv[5]));
		if (argc == 7 && !std::strcmp(argv[1], "rateone"))
			return run_rate_one(be, argv[2], argv[3], std::atof(argv[4]), std::atof(argv[5]),
			                    (std::size_t)std::atol(argv[6]));
		if (argc == 6 && !std::strcmp(argv[1], "explicit"))
			return run_explicit(be, argv[2], std::atof(argv[3]), std::atof(argv[4]),
			                    (std::size_t)std::atol(argv[5]));
		if (argc == 6 && !std::strcmp(argv[1], "heat"))
			return run_heat(be, argv[2], std::atoi(argv[3]), std::atof(argv[4]), argv[5]);
		std::fprintf(stderr, "usage: bdf_check rate <cfg> <lambda> <ic> <n>\n"
		                     "       bdf_check rateone <cfg> <section> <lambda> <ic> <n>\n"
		                     "       bdf_check explicit <cfg> <lambda> <ic> <n>\n"
		                     "       bdf_check heat <cfg> <n> <alpha> <out.bin>\n");
		return 2;
	}
	catch (const std::exception & e) {
		std::fprintf(stderr, "error: %s\n", e.what());
		return 1;
	}
}
