// flecsolve/time-integrators/rk45.hh -- explicit Runge-Kutta-Fehlberg 4(5)
// with embedded error control.
//
// Drop-in for the reference's flecsolve/time-integrators/rk45.hh:23-148:
// the settings, options and parameters of rk23 (rk45.hh:25-40) and the six
// Fehlberg stages (coefficients 1/4; 3/32, 9/32; 1932/2197, -7200/2197,
// 7296/2197; 439/216, -8, 3680/513, -845/4104; -8/27, 2, -3544/2565,
// 1859/4104, -11/40), each stage input formed by the reference's chain of
// axpy operations in its order; the fourth-order combination (25/216,
// 1408/2565, 2197/4104, -1/5) goes to z, the fifth-order one (16/135,
// 6656/12825, 28561/56430, -9/50, 2/55) to the new state, and z = u5 - u4 is
// the error estimate.  A rejected step shrinks dt by the safety factor; an
// accepted one takes safety * dt * (atol / |z|)^(1/5), clamped.
#ifndef FLECSOLVE_AMD_TIME_INTEGRATORS_RK45_HH
#define FLECSOLVE_AMD_TIME_INTEGRATORS_RK45_HH

#include "flecsolve/time-integrators/rk23.hh"

namespace flecsolve::time_integrator::rk45 {

struct settings : rk23::settings {};

struct options : rk23::options {
	using settings_type = settings;
	explicit options(const char * pre) : rk23::options(pre) {}
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

enum workvecs : std::size_t { k1, k2, k3, k4, k5, k6, z, next, nvecs };

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
		auto & [k1, k2, k3, k4, k5, k6, z, next] = params.work;
		F.apply(curr, k1);
		next.axpy(0.25 * dt, k1, curr);
		F.apply(next, k2);
		next.axpy(3.0 * dt / 32., k1, curr);
		next.axpy(9.0 * dt / 32., k2, next);
		F.apply(next, k3);
		next.axpy(1932. * dt / 2197., k1, curr);
		next.axpy(-7200. * dt / 2197., k2, next);
		next.axpy(7296. * dt / 2197., k3, next);
		F.apply(next, k4);
		next.axpy(439. * dt / 216., k1, curr);
		next.axpy(-8. * dt, k2, next);
		next.axpy(3680. * dt / 513., k3, next);
		next.axpy(-845. * dt / 4104., k4, next);
		F.apply(next, k5);
		next.axpy(-8. * dt / 27., k1, curr);
		next.axpy(2. * dt, k2, next);
		next.axpy(-3544. * dt / 2565., k3, next);
		next.axpy(1859. * dt / 4104., k4, next);
		next.axpy(-11. * dt / 40., k5, next);
		F.apply(next, k6);
		z.axpy(25. * dt / 216., k1, curr);
		z.axpy(1408. * dt / 2565., k3, z);
		z.axpy(2197. * dt / 4104., k4, z);
		z.axpy(-0.2 * dt, k5, z);
		next.axpy(16. * dt / 135., k1, curr);
		next.axpy(6656. * dt / 12825., k3, next);
		next.axpy(28561. * dt / 56430., k4, next);
		next.axpy(-9. * dt / 50., k5, next);
		next.axpy(2. * dt / 55., k6, next);
		z.subtract(next, z);
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
		else if (good) {
			const double err = std::get<workvecs::z>(params.work).l2norm().get();
			next_dt = params.safety_factor * current_dt * std::pow((params.atol / err), 1. / 5.);
			next_dt = std::min(std::max(next_dt, params.min_dt), params.max_dt);
			next_dt = std::min(next_dt, params.final_time - current_time);
		}
		else {
			next_dt = params.safety_factor * current_dt;
			++total_step_rejects;
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
