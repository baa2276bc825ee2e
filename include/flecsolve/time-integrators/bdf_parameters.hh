// flecsolve/time-integrators/bdf_parameters.hh -- settings, config options
// and parameters of the BDF / Crank-Nicolson integrator.
//
// Same enums, fields, config keys, requirements and defaults as the
// reference's flecsolve/time-integrators/bdf_parameters.hh:29-171 and
// bdf_parameters.cc (token spellings of the stream extractions), and the same
// validate() overrides (bdf_parameters.hh:62-107), with failed checks thrown
// as std::invalid_argument.  Options the reference declares without a
// default (calculate-time-trunc-error, time-trunc-error-norm,
// target-relative-change, problem-scales) keep the member initialisers below
// when the file omits them.  memory_size and order follow bdf.cc:5-22.
#ifndef FLECSOLVE_AMD_TIME_INTEGRATORS_BDF_PARAMETERS_HH
#define FLECSOLVE_AMD_TIME_INTEGRATORS_BDF_PARAMETERS_HH

#include <algorithm>
#include <cctype>
#include <istream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "flecsolve/time-integrators/parameters.hh"
#include "flecsolve/util/config.hh"
#include "flecsolve/vectors/util.hh"

namespace flecsolve::time_integrator::bdf {

enum class method { cn, be, bdf2, bdf3, bdf4, bdf5, bdf6 };
enum class predictor { ab2, leapfrog };
enum class strategy { truncation_error, constant, final_constant, limit_relative_change };
enum class controller { H211b, pc4_7, pc11, deadbeat };
enum class error_scaling { fixed_resolution, fixed_scaling };

// number of previous solutions a method uses: 1 for CN and BE, k for BDFk
inline int memory_size(method m) {
	const int v = static_cast<int>(m);
	return v == 0 ? 1 : v;
}

inline short order(method m) {
	switch (m) {
	case method::cn: return 2;
	case method::be: return 1;
	default: return static_cast<short>(m);
	}
}

namespace detail {
inline std::string token(std::istream & in, bool lower) {
	std::string t;
	in >> t;
	if (lower)
		std::transform(t.begin(), t.end(), t.begin(),
		               [](unsigned char c) { return (char)std::tolower(c); });
	return t;
}
}

inline std::istream & operator>>(std::istream & in, predictor & p) {
	const auto t = detail::token(in, true);
	if (t == "ab2") p = predictor::ab2;
	else if (t == "leapfrog") p = predictor::leapfrog;
	else in.setstate(std::ios_base::failbit);
	return in;
}

inline std::istream & operator>>(std::istream & in, strategy & s) {
	const auto t = detail::token(in, true);
	if (t == "truncation-error") s = strategy::truncation_error;
	else if (t == "constant") s = strategy::constant;
	else if (t == "final-constant") s = strategy::final_constant;
	else if (t == "limit-relative-change") s = strategy::limit_relative_change;
	else in.setstate(std::ios_base::failbit);
	return in;
}

inline std::istream & operator>>(std::istream & in, method & m) {
	const auto t = detail::token(in, false);
	if (t == "BE") m = method::be;
	else if (t == "CN") m = method::cn;
	else if (t == "BDF2") m = method::bdf2;
	else if (t == "BDF3") m = method::bdf3;
	else if (t == "BDF4") m = method::bdf4;
	else if (t == "BDF5") m = method::bdf5;
	else if (t == "BDF6") m = method::bdf6;
	else in.setstate(std::ios_base::failbit);
	return in;
}

inline std::istream & operator>>(std::istream & in, controller & c) {
	const auto t = detail::token(in, false);
	if (t == "PC.4.7") c = controller::pc4_7;
	else if (t == "H211b") c = controller::H211b;
	else if (t == "PC11") c = controller::pc11;
	else if (t == "Deadbeat") c = controller::deadbeat;
	else in.setstate(std::ios_base::failbit);
	return in;
}

inline std::istream & operator>>(std::istream & in, error_scaling & s) {
	const auto t = detail::token(in, false);
	if (t == "fixed-scaling") s = error_scaling::fixed_scaling;
	else if (t == "fixed-resolution") s = error_scaling::fixed_resolution;
	else in.setstate(std::ios_base::failbit);
	return in;
}

struct settings : base_settings {
	bool use_predictor = true;
	bool use_initial_predictor = true;
	bool has_source_term = false;
	bool combine_timestep_estimators = false;
	bdf::predictor predictor = bdf::predictor::leapfrog;
	bdf::strategy timestep_strategy = bdf::strategy::truncation_error;
	double dt_cut_lower_bound = 0.58754407;
	double dt_growth_upper_bound = 1.702;
	int number_of_time_intervals = 100;
	bdf::method integrator = bdf::method::bdf2;
	bdf::method starting_integrator = bdf::method::cn;
	bool calculate_time_trunc_error = false;
	vec::norm_type time_trunc_err_norm = vec::norm_type::inf;
	double target_relative_change = 0;
	bool use_pi_controller = true;
	bdf::controller pi_controller_type = bdf::controller::pc4_7;
	bool control_timestep_variation = false;
	bdf::error_scaling time_error_scaling = bdf::error_scaling::fixed_scaling;
	double time_rtol = 1e-9, time_atol = 1e-15;
	std::vector<double> problem_scales;

	void validate() {
		auto require = [](bool c, const char * what) {
			if (!c) throw std::invalid_argument(std::string("BDF settings: ") + what);
		};
		require(memory_size(starting_integrator) == 1, "Starting integrator must be CN or BE");
		if (use_predictor) calculate_time_trunc_error = true;
		if (timestep_strategy == strategy::truncation_error) {
			calculate_time_trunc_error = true;
			use_predictor = true;
		}
		else {
			combine_timestep_estimators = false;
			control_timestep_variation = false;
			use_pi_controller = false;
		}
		if (calculate_time_trunc_error) {
			use_predictor = true;
			if (time_error_scaling == error_scaling::fixed_scaling)
				require(!problem_scales.empty(),
				        "Problem scales must be specified if using fixed time error scaling");
		}
		if (use_predictor) {
			if (memory_size(integrator) > 1)
				require(predictor == bdf::predictor::leapfrog,
				        "Only the leapfrog predictor is supported for BDF2-6");
			else if (integrator == method::cn)
				require(predictor == bdf::predictor::ab2,
				        "Valid option for Crank-Nicolson predictor is only ab2 currently");
		}
	}
};

struct options : base_options {
	using settings_type = settings;
	explicit options(const char * pre) : base_options(pre) {}

	optdesc operator()(settings & s) {
		optdesc d = base_options::operator()(s);
		d.with_default(label("use-predictor"), &s.use_predictor, true)
			.with_default(label("has-source-term"), &s.has_source_term, false)
			.with_default(label("use-initial-predictor"), &s.use_initial_predictor, true)
			.required(label("predictor"), &s.predictor)
			.required(label("timestep-selection-strategy"), &s.timestep_strategy)
			.with_default(label("dt-cut-lower-bound"), &s.dt_cut_lower_bound, 0.58754407)
			.with_default(label("dt-growth-upper-bound"), &s.dt_growth_upper_bound, 1.702)
			.with_default(label("number-of-time-intervals"), &s.number_of_time_intervals, 100)
			.required(label("integrator"), &s.integrator)
			.required<method>(label("starting-integrator"), &s.starting_integrator,
			                  [](const method & m) {
				                  if (memory_size(m) != 1)
					                  throw std::invalid_argument(
						                  "BE or CN must be used for starting integrator");
			                  })
			.optional(label("calculate-time-trunc-error"), &s.calculate_time_trunc_error)
			.optional(label("time-trunc-error-norm"), &s.time_trunc_err_norm)
			.optional(label("target-relative-change"), &s.target_relative_change)
			.with_default(label("use-pi-controller"), &s.use_pi_controller, true)
			.with_default(label("pi-controller-type"), &s.pi_controller_type,
			              controller::pc4_7)
			.with_default(label("control-timestep-variation"), &s.control_timestep_variation,
			              false)
			.with_default(label("time-error-scaling"), &s.time_error_scaling,
			              error_scaling::fixed_scaling)
			.with_default(label("trunc-error-rtol"), &s.time_rtol, 1e-9)
			.with_default(label("trunc-error-atol"), &s.time_atol, 1e-15)
			.with_default(label("combine-timestep-estimators"), &s.combine_timestep_estimators,
			              false)
			.optional(label("problem-scales"), &s.problem_scales);
		return d;
	}
};

// settings, right-hand-side operator, work vectors and the linear solver of
// the stage system (bdf_parameters.hh:150-166)
template<class Op, class Work, class Solver>
struct parameters : time_integrator::parameters<settings, Op, Work> {
	using base = time_integrator::parameters<settings, Op, Work>;

	template<class W>
	parameters(const settings & s, op::handle<Op> op, W && work, op::handle<Solver> solver)
		: base(s, op, std::forward<W>(work)), solver(solver) {}

	auto & get_solver() { return solver.get(); }

protected:
	op::handle<Solver> solver;
};

template<class O, class W, class S>
parameters(const settings &, op::handle<O>, W &&, op::handle<S>) -> parameters<O, W, S>;

}

#endif
