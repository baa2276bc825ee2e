// flecsolve/time-integrators/bdf.hh -- variable-step BDF1-6 / Crank-Nicolson
// integrator with predictors and local-truncation-error step control.
//
// Drop-in for the reference's flecsolve/time-integrators/bdf.hh:51-947:
//
//   bdf::integrator ti(bdf::parameters(settings, F, bdf::make_work(u), solver));
//   while (ti.get_current_time() < ti.get_final_time()) {
//       ti.advance(dt, first_step, u, unew);      // solve one stage
//       bool good = ti.check_solution();          // LTE test
//       if (good) { ti.update(); std::swap(u, unew); first_step = false; }
//       dt = ti.get_next_dt(good);
//   }
//
// F is a handle to the right-hand-side operator (apply_rhs(x, y) gives
// y = f(x), set_scaling(gamma) sets the stage operator x - gamma f(x),
// is_valid(v)); solver is a handle whose apply(b, x) solves the stage system
// from the guess in x and returns a solve_info.  Every step is the
// reference's sequence of vector operations in the reference's order --
// predictor, source term, stage solve, time-derivative estimate, scaled LTE
// norm, PI / deadbeat step control, including its side effects (a
// forward-Euler predictor evaluated while first_step is still set refreshes
// the stored time derivative; the LTE history carries over rejected steps;
// the final-constant strategy counts in a function-local static shared by all
// integrators of one type; BDF6's sixth coefficient divides by h1 - h5 where
// the textbook formula has h1 - h6) -- so that a run takes the reference's
// steps and rejections.  On device vectors each operation is one libfsx
// kernel and each norm one device reduction read back to the host.
//
// Members the reference leaves uninitialised start as: current step rejects
// 0, previous successive rejects false, truncation error estimate 1 (the value
// it assigns when it has no estimate).
#ifndef FLECSOLVE_AMD_TIME_INTEGRATORS_BDF_HH
#define FLECSOLVE_AMD_TIME_INTEGRATORS_BDF_HH

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <functional>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <utility>
#include <vector>

#include "flecsolve/time-integrators/base.hh"
#include "flecsolve/time-integrators/bdf_parameters.hh"
#include "flecsolve/vectors/util.hh"

namespace flecsolve::time_integrator::bdf {

enum workvec : std::size_t {
	rhs,
	source,
	current_sol,
	scratch,
	previous_function,
	current_function,
	scratch_function,
	old_td,
	predict,
	time_deriv,
	nwork
};

// the ten named vectors and up to six previous solutions (BDF6)
inline constexpr std::size_t work_size = workvec::nwork + 6;

// make_work(u): work vectors shaped like u
struct work_maker {
	template<class V>
	auto operator()(const V & v) const { return vec::make_work<work_size>(v); }
};
inline constexpr work_maker make_work{};

template<class O, class W, class S>
struct integrator : base<parameters<O, W, S>> {
	using P = parameters<O, W, S>;
	using base<P>::params;
	using base<P>::current_dt;
	using base<P>::old_dt;
	using base<P>::current_time;
	using base<P>::assert_can_advance;
	using base<P>::integrator_step;

	explicit integrator(P p)
		: base<P>(std::move(p)), prev(memory_size(params.integrator)) {
		params.validate();
		getvec<workvec::scratch>().zero();
	}

	// bdf.hh:74-95
	template<class Curr, class Out>
	void advance(double dt, bool first, Curr & curr, Out & out) {
		if (curr == out)
			throw std::invalid_argument("BDF integrator: curr cannot be the same as out");
		assert_can_advance();
		hist(0).copy(curr);
		current_dt = dt;
		set_initial_guess(first);
		auto & r = getvec<workvec::rhs>();
		auto & sol = getvec<workvec::current_sol>();
		r.scale(-1., getvec<workvec::source>());
		auto info = params.get_solver().apply(r, sol);
		solver_success = info.success();
		out.copy(sol);
	}

	double get_time_operator_scaling() const { return gamma; }

	// bdf.hh:99-130
	bool check_solution() {
		bool passed = false;
		if (solver_success) {
			if (params.calculate_time_trunc_error) calculate_temporal_truncation_error();
			if (params.timestep_strategy == strategy::truncation_error) {
				passed = time_trunc_err_est <= 1.0;
				if (passed) {
					prev_successive_rejects = current_steprejects > 1;
					current_steprejects = 0;
				}
				else
					++current_steprejects;
			}
			else
				passed = true;
		}
		else
			++current_steprejects;
		if (!passed) ++total_steprejects;
		return passed;
	}

	// bdf.hh:132-149
	void update() {
		current_time += current_dt;
		new_time = current_time;
		if (params.use_predictor) {
			if (params.predictor == predictor::ab2)
				std::swap(getvec<workvec::old_td>(), getvec<workvec::current_function>());
			estimate_time_derivative();
		}
		push(getvec<workvec::current_sol>(), current_dt);
		++integrator_step;
	}

	// bdf.hh:151-219
	double get_next_dt(bool good) {
		const double tmp_dt = current_dt;
		if (params.timestep_strategy == strategy::truncation_error) {
			if (good) {
				current_dt = estimate_dt_with_truncation_error_estimates(current_dt, good);
				if (params.max_dt < current_dt) current_dt = params.max_dt;
			}
			else if (solver_success)
				current_dt = estimate_dt_with_truncation_error_estimates(current_dt, good);
			else
				current_dt = params.dt_cut_lower_bound * current_dt;
			evaluate_predictor();
			auto & op = params.get_operator();
			if (!op.is_valid(getvec<workvec::predict>())) {
				for (int i = 0; i < 10; ++i) {
					current_dt = 0.5 * current_dt;
					evaluate_predictor();
					op.is_valid(getvec<workvec::predict>());
				}
			}
		}
		else if (good) {
			if (params.timestep_strategy == strategy::constant)
				current_dt = params.initial_dt;
			else if (params.timestep_strategy == strategy::final_constant) {
				static int i = 1;
				if (i < params.number_of_time_intervals) {
					current_dt = params.initial_dt +
					             ((double)i) / ((double)params.number_of_time_intervals) *
					                 (params.max_dt - params.initial_dt);
					++i;
				}
				else
					current_dt = params.max_dt;
			}
			else if (params.timestep_strategy == strategy::limit_relative_change)
				current_dt = estimate_dynamical_time_scale(current_dt);
		}
		else
			current_dt = 0.;
		if (good) old_dt = tmp_dt;
		current_dt = std::min(std::min(current_dt, params.max_dt),
		                      params.final_time - current_time);
		return current_dt;
	}

	int num_step_rejects() const { return total_steprejects; }

protected:
	using vec_t = typename std::decay_t<W>::value_type;

	// ring of memory_size(integrator) previous solutions and their steps
	// (bdf.hh:880-947): slot k of the ring is work vector nwork + k; entry 0 is
	// the newest.
	struct history {
		explicit history(std::size_t len) : dts(len, 0.0), pos(0), len(len) {}
		std::size_t slot(std::size_t i) const {
			const std::size_t beg = len - 1 - pos;
			return len - 1 - (beg + i + 1) % len;
		}
		std::vector<double> dts;
		std::size_t pos, len;
	};

	vec_t & hist(std::size_t i) { return params.work[workvec::nwork + prev.slot(i)]; }
	double hist_dt(std::size_t i) const { return prev.dts[prev.slot(i)]; }
	void push(const vec_t & v, double dt) {
		params.work[workvec::nwork + prev.pos].copy(v);
		prev.dts[prev.pos] = dt;
		prev.pos = (prev.pos + 1) % prev.len;
	}

	template<workvec w>
	vec_t & getvec() { return std::get<w>(params.work); }

	// bdf.hh:224-234
	method get_current_integrator() const {
		if (integrator_step == 0) return params.starting_integrator;
		if (integrator_step + 1 < memory_size(params.integrator))
			return static_cast<method>(integrator_step + 1);
		return params.integrator;
	}

	// bdf.hh:236-277
	void set_initial_guess(bool first) {
		first_step = first;
		if (get_current_integrator() == method::cn) {
			auto & sc = getvec<workvec::scratch>();
			sc.copy(hist(0));
			params.get_operator().apply_rhs(sc, getvec<workvec::previous_function>());
		}
		auto & sol = getvec<workvec::current_sol>();
		if (params.use_predictor) {
			auto & pr = getvec<workvec::predict>();
			if (first && !params.use_initial_predictor) pr.copy(hist(0));
			else evaluate_predictor();
			sol.copy(pr);
		}
		else
			sol.copy(hist(0));
		compute_source_term();
	}

	// BDF2-6 coefficients a1..ak of the previous solutions and a0 = gamma for
	// h[i] = t_{n+1} - t_{n-i} (bdf.hh:309-437); k = memory_size(m).  The
	// member alpha takes the sum of the products n_i, as in the reference.
	void bdf_coefficients(method m, const std::vector<double> & h, double * a, double & a0) {
		const int k = memory_size(m);
		// n_i: product of the h_j, j != i (the reference forms each product
		// starting at h_{i+1} and wrapping round)
		double nn[6] = {};
		for (int i = 0; i < k; ++i) {
			double p = h[(i + 1) % k];
			for (int j = 2; j < k; ++j) p *= h[(i + j) % k];
			nn[i] = p;
		}
		double al = nn[0];
		for (int i = 1; i < k; ++i) al += nn[i];
		alpha = al;
		double hp = h[0];
		for (int i = 1; i < k; ++i) hp *= h[i];
		a0 = hp / al;
		for (int i = 0; i < k; ++i) {
			double den = al;
			for (int j = 0; j < k; ++j) {
				if (j == i) continue;
				// den = alpha * prod_{j != i} (h_j - h_i), in increasing j; the
				// reference's BDF6 a6 takes h1 - h5 as its first factor
				const int jj = (m == method::bdf6 && i == 5 && j == 0) ? 4 : i;
				den *= (h[j] - h[jj]);
			}
			a[i] = -(nn[i] * nn[i]) / den;
		}
	}

	// bdf.hh:279-444
	void compute_source_term() {
		auto & f = getvec<workvec::source>();
		const auto m = get_current_integrator();
		std::vector<double> h(memory_size(m));
		h[0] = current_dt;
		for (int i = 1; i < memory_size(m); ++i) h[i] = h[i - 1] + hist_dt(i - 1);
		switch (m) {
		case method::be:
			f.scale(-1., hist(0));
			gamma = current_dt;
			break;
		case method::cn: {
			const double eps = 0.0;
			const double a = 0.5 * current_dt + eps;
			const double b = 0.5 * current_dt - eps;
			f.axpy(b, getvec<workvec::previous_function>(), hist(0));
			f.scale(-1.);
			gamma = a;
			break;
		}
		default: {
			double a[6] = {}, a0 = 0;
			bdf_coefficients(m, h, a, a0);
			f.scale(a[0], hist(0));
			for (int i = 1; i < memory_size(m); ++i) f.axpy(a[i], hist(i), f);
			gamma = a0;
		}
		}
		params.get_operator().set_scaling(gamma);
	}

	// bdf.hh:450-505 (Gresho and Sani estimates)
	void estimate_time_derivative() {
		auto & td = getvec<workvec::time_deriv>();
		auto & cs = getvec<workvec::current_sol>();
		switch (get_current_integrator()) {
		case method::cn:
			td.axpy(-1., hist(0), cs);
			td.linear_sum(2. / current_dt, td, -1., getvec<workvec::previous_function>());
			break;
		case method::be:
			td.axpy(-1., hist(0), cs);
			td.scale(1. / current_dt);
			break;
		default: {
			const double dtt = current_dt / old_dt;
			const double a = (2. * dtt + 1.) / ((dtt + 1.) * current_dt);
			const double b = -(dtt + 1.) / current_dt;
			const double g = (dtt * dtt) / ((dtt + 1.) * current_dt);
			td.linear_sum(a, cs, b, hist(0));
			td.axpy(g, hist(1), td);
		}
		}
	}

	// bdf.hh:507-546
	void evaluate_predictor() {
		switch (params.integrator) {
		case method::cn:
			if (params.predictor != predictor::ab2)
				throw std::logic_error("Crank-Nicolson currently only supports ab2 predictor");
			evaluate_ab2_predictor();
			break;
		case method::be:
			evaluate_forward_euler_predictor();
			break;
		default:
			if (first_step) {
				if (params.starting_integrator != method::cn &&
				    params.starting_integrator != method::be)
					throw std::logic_error("starting integrator must be CN or BE");
				evaluate_forward_euler_predictor();
			}
			else {
				if (params.predictor != predictor::leapfrog)
					throw std::logic_error("BDF2-6 currently only supports the leapfrog predictor");
				evaluate_leapfrog_predictor();
			}
		}
		if (!params.get_operator().is_valid(getvec<workvec::predict>()))
			evaluate_forward_euler_predictor();
	}

	// bdf.hh:548-560
	void evaluate_ab2_predictor() {
		const double r = current_dt / old_dt;
		const double a = current_dt * (2. + r) / 2.;
		const double b = -current_dt * r / 2.;
		auto & pr = getvec<workvec::predict>();
		pr.linear_sum(a, getvec<workvec::current_function>(), b, getvec<workvec::old_td>());
		pr.add(pr, hist(0));
	}

	// bdf.hh:562-573
	void evaluate_leapfrog_predictor() {
		const double r = current_dt / old_dt;
		const double a = r * r;
		const double b = 1. - a;
		const double g = (1. + r) * current_dt;
		auto & pr = getvec<workvec::predict>();
		pr.linear_sum(a, hist(1), b, hist(0));
		pr.axpy(g, getvec<workvec::time_deriv>(), pr);
	}

	// bdf.hh:575-595
	void evaluate_forward_euler_predictor() {
		const double g = current_dt;
		auto & td = getvec<workvec::time_deriv>();
		if (first_step) {
			auto & sc = getvec<workvec::scratch>();
			auto & cf = getvec<workvec::current_function>();
			sc.copy(hist(0));
			params.get_operator().apply_rhs(sc, cf);
			td.copy(cf);
		}
		auto & pr = getvec<workvec::predict>();
		pr.copy(hist(0));
		pr.axpy(g, td, pr);
	}

	double lte_norm(const vec_t & v) const {
		return params.time_trunc_err_norm == vec::norm_type::inf ? (double)v.inf_norm().get()
		                                                          : (double)v.l2norm().get();
	}

	// bdf.hh:601-644: the relative change is formed in scratch, and the norm
	// is then taken of scratch_function, as the reference does
	double estimate_dynamical_time_scale(double curr_dt) {
		if (params.integrator == method::cn)
			std::cerr << "BDF integrator: relative-change control is not implemented for CN\n";
		auto & sc = getvec<workvec::scratch>();
		sc.add_scalar(hist(0), std::numeric_limits<double>::epsilon());
		sc.divide(hist(1), sc);
		sc.add_scalar(sc, -1);
		sc.abs(sc);
		const double change = lte_norm(getvec<workvec::scratch_function>());
		double factor = (curr_dt > 0.5) ? 1.05 : 1.1;
		factor = (curr_dt > 10.) ? 1.03 : factor;
		const double cfl_new_dt = std::sqrt(params.target_relative_change / change) * curr_dt;
		return std::min(cfl_new_dt, factor * curr_dt);
	}

	// bdf.hh:650-727
	double estimate_dt_with_truncation_error_estimates(double dt, bool good) {
		double f = 0.;
		const double p = static_cast<double>(order(get_current_integrator()));
		const double eps = std::numeric_limits<double>::epsilon();
		if (params.use_pi_controller && good && !prev_successive_rejects) {
			const double sf = 0.8;
			switch (params.pi_controller_type) {
			case controller::H211b: {
				const double b = 4.;
				f = std::pow(sf / std::max(time_trunc_err_est, eps), 1. / (b * (p + 1.)));
				f *= std::pow(sf / std::max(prev_time_trunc_err_est, eps), 1. / (b * (p + 1.)));
				f *= std::pow(alpha, -0.25);
				break;
			}
			case controller::pc4_7:
				f = std::pow(sf / std::max(time_trunc_err_est, eps), 0.4 / (p + 1.));
				f *= std::pow(time_err_est_ratio, 0.7 / (p + 1.)) * alpha;
				break;
			case controller::pc11:
				f = std::pow(sf / std::max(time_trunc_err_est, eps), 1.0 / (p + 1.));
				f *= std::pow(time_err_est_ratio, 1. / (p + 1.)) * alpha;
				break;
			case controller::deadbeat:
				f = std::pow(sf / std::max(time_trunc_err_est, eps), 1. / (p + 1.));
				break;
			}
		}
		else {
			const double sf = 0.8;
			f = std::pow(sf / std::max(time_trunc_err_est, eps), 1. / (p + 1.));
		}
		if (params.control_timestep_variation && !params.use_pi_controller)
			if (std::fabs(1. - f) < 0.1) f = 1.;
		current_dt = dt * f;
		return current_dt;
	}

	// bdf.hh:729-763
	double calculate_LTE_scaling_factor() const {
		double ef = 0.;
		if (first_step) {
			if (params.starting_integrator == method::be) ef = 0.5;
			else if (params.starting_integrator == method::cn) {
				ef = 1.5;
				ef *= current_dt;
			}
		}
		else {
			switch (params.predictor) {
			case predictor::leapfrog: ef = (1. + alpha) / (2. + 3. * alpha); break;
			case predictor::ab2: ef = 2. / (6. - 1. / ((alpha + 1) * (alpha + 1))); break;
			}
		}
		return ef;
	}

	// bdf.hh:765-802 (single-component vectors: one norm)
	double calculate_scaled_LTE_norm(const vec_t & x, const vec_t & y) {
		auto & sc = getvec<workvec::scratch>();
		switch (params.time_error_scaling) {
		case error_scaling::fixed_resolution:
			sc.scale(params.time_rtol, x);
			sc.add_scalar(sc, params.time_atol);
			break;
		case error_scaling::fixed_scaling:
			sc.zero();
			sc.add_scalar(x, params.problem_scales[0]);
			sc.scale(params.time_rtol);
			break;
		}
		auto & sf = getvec<workvec::scratch_function>();
		sf.subtract(x, y);
		sf.divide(sf, sc);
		return lte_norm(sf);
	}

	// bdf.hh:804-855
	void calculate_temporal_truncation_error() {
		auto & sol = getvec<workvec::current_sol>();
		if (integrator_step > 0 || (first_step && params.use_initial_predictor)) {
			alpha = current_dt / old_dt;
			const double ef = calculate_LTE_scaling_factor();
			auto & y = (first_step && params.starting_integrator == method::cn)
			               ? hist(0)
			               : getvec<workvec::predict>();
			const double est = calculate_scaled_LTE_norm(sol, y) * ef;
			prev_time_trunc_err_est = time_trunc_err_est;
			time_trunc_err_est = std::max(std::numeric_limits<double>::epsilon(), est);
			time_err_est_ratio = prev_time_trunc_err_est / time_trunc_err_est;
		}
		else {
			prev_time_trunc_err_est = 1.;
			time_trunc_err_est = 1.;
			time_err_est_ratio = 1.;
		}
	}

	bool prev_successive_rejects = false;
	int current_steprejects = 0;
	double new_time = 0;
	double time_trunc_err_est = 1.;
	double prev_time_trunc_err_est = 1.;
	double alpha = 1.;
	double time_err_est_ratio = 1.;
	bool first_step = true;
	double gamma = 0;
	history prev;
	bool solver_success = false;
	int total_steprejects = 0;
};

template<class O, class W, class S>
integrator(parameters<O, W, S>) -> integrator<O, W, S>;

}

#endif
