// flecsolve/time-integrators/rk23.hh -- explicit Runge-Kutta 2(3)
// (Bogacki-Shampine) with embedded error control.
//
// Drop-in for the reference's flecsolve/time-integrators/rk23.hh:23-143:
// settings (safety-factor 0.9, atol 1e-9 -- both float, as in the reference
// -- and use-fixed-dt false by default, after the base keys), parameters
// (settings, operator handle, work vectors) and the integrator with
// advance(dt, curr, out), check_solution(), update(), get_next_dt(good).
// The operator's apply(x, y) is y = f(x).  advance forms the stages
//   k1 = f(u), k2 = f(u + dt/2 k1), k3 = f(u + 3dt/4 k2),
//   u+ = u + dt/9 (2 k1 + 3 k2 + 4 k3), k4 = f(u+),
//   z = dt/72 (-5 k1 + 6 k2 + 8 k3 - 9 k4)       (error estimate)
// with the reference's vector operations in its order; a step is accepted
// when |z|_2 < atol or dt is within 1e-10 of min-dt; the next step is
// safety * dt * (atol / |z|)^(1/3) clamped to [min-dt, max-dt] and the time
// left (or the same dt with use-fixed-dt).
#ifndef FLECSOLVE_AMD_TIME_INTEGRATORS_RK23_HH
#define FLECSOLVE_AMD_TIME_INTEGRATORS_RK23_HH

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <tuple>
#include <utility>

#include "flecsolve/time-integrators/base.hh"
#include "flecsolve/time-integrators/parameters.hh"
#include "flecsolve/vectors/util.hh"

namespace flecsolve::time_integrator::rk23 {

struct settings : base_settings {
	float safety_factor = 0.9f;
	float atol = 1e-9f;
	bool use_fixed_dt = false;
};

struct options : base_options {
	using settings_type = settings;
	explicit options(const char * pre) : base_options(pre) {}

	optdesc operator()(settings & s) {
		optdesc d = base_options::operator()(s);
		d.with_default(label("safety-factor"), &s.safety_factor, 0.9f)
			.with_default(label("atol"), &s.atol, 1e-9f)
			.with_default(label("use-fixed-dt"), &s.use_fixed_dt, false);
		return d;
	}
};

template<class Op, class Work>
struct parameters : time_integrator::parameters<settings, Op, Work> {
	using base = time_integrator::parameters<settings, Op, Work>;
	template<class W>
	parameters(const settings & s, op::handle<Op> op, W && work)
		: base(s, op, std::forward<W>(work)) {}
};
template<class O, class W>
parameters(const settings &, op::handle<O>, W &&) -> parameters<O, W>;

enum workvecs : std::size_t { k1, k2, k3, k4, z, next, nvecs };

struct work_maker {
	template<class V>
	auto operator()(const V & v) const { return vec::make_work<workvecs::nvecs>(v); }
};
inline constexpr work_maker make_work{};

template<class O, class W>
struct integrator : base<parameters<O, W>> {
	using P = parameters<O, W>;
	using base<P>::params;
	using base<P>::current_dt;
	using base<P>::current_time;
	using base<P>::assert_can_advance;

	explicit integrator(P p) : base<P>(std::move(p)) {}

	template<class Curr, class Out>
	void advance(double dt, Curr & curr, Out & out) {
		assert_can_advance();
		current_dt = dt;
		auto & F = params.get_operator();
		auto & [k1, k2, k3, k4, z, next] = params.work;
		F.apply(curr, k1);
		next.axpy(0.5 * dt, k1, curr);
		F.apply(next, k2);
		next.axpy(0.75 * dt, k2, curr);
		F.apply(next, k3);
		next.linear_sum(2.0, k1, 3.0, k2);
		next.axpy(4.0, k3, next);
		next.axpy(dt / 9.0, next, curr);
		F.apply(next, k4);
		z.linear_sum(-5., k1, 6., k2);
		z.axpy(8., k3, z);
		z.axpy(-9., k4, z);
		z.scale(dt / 72.);
		out.copy(next);
	}

	bool check_solution() {
		const double err = std::get<workvecs::z>(params.work).l2norm().get();
		return err < params.atol || std::fabs(current_dt - params.min_dt) < 1e-10;
	}

	double get_next_dt(bool good) {
		double next_dt;
		if (params.use_fixed_dt)
			next_dt = std::min(current_dt, params.final_time - current_time);
		else {
			const double err = std::get<workvecs::z>(params.work).l2norm().get();
			next_dt = params.safety_factor * current_dt * std::pow(params.atol / err, 1. / 3.);
			next_dt = std::min(std::max(next_dt, params.min_dt), params.max_dt);
			next_dt = std::min(next_dt, params.final_time - current_time);
			if (!good) ++total_step_rejects;
		}
		return next_dt;
	}

	int num_step_rejects() const { return total_step_rejects; }

protected:
	int total_step_rejects = 0;
};
template<class O, class W>
integrator(parameters<O, W>) -> integrator<O, W>;

}

#endif
