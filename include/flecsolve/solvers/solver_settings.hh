// flecsolve/solvers/solver_settings.hh -- solver settings and results.
//
// Same fields and types as the reference's
// flecsolve/solvers/solver_settings.hh:31-85: rtol and atol are float (so a
// configured 1e-10 is used as (double)1.00000001e-10f), solve_info norms are
// float, and success() accepts the three converged reasons.  solver_options
// (solver_settings.hh:36-56) declares the same config-file keys for
// read_config (flecsolve/util/config.hh).
#ifndef FLECSOLVE_AMD_SOLVERS_SOLVER_SETTINGS_HH
#define FLECSOLVE_AMD_SOLVERS_SOLVER_SETTINGS_HH

#include <ostream>

#include "flecsolve/util/config.hh"

namespace flecsolve {

struct solver_settings {
	int maxiter = 100;
	float rtol = 0;
	float atol = 0;
	bool use_zero_guess = false;
};

// maxiter and use-zero-guess required, rtol and atol default 0
struct solver_options : with_label {
	using settings_type = solver_settings;
	solver_options(const char * pre) : with_label(pre) {}
	optdesc operator()(solver_settings & s) {
		optdesc d;
		d.required(label("maxiter"), &s.maxiter)
			.with_default(label("rtol"), &s.rtol, 0.0f)
			.with_default(label("atol"), &s.atol, 0.0f)
			.required(label("use-zero-guess"), &s.use_zero_guess);
		return d;
	}
	const std::string & get_prefix() const { return prefix; }
};

struct solve_info {
	enum class stop_reason {
		converged_atol,
		converged_rtol,
		converged_user,
		diverged_dtol,
		diverged_iters,
		diverged_breakdown,
		unknown
	};
	stop_reason status = stop_reason::unknown;
	int iters = 0;
	int restarts = 0;
	float res_norm_initial = 0, res_norm_final = 0;
	float sol_norm_initial = 0, sol_norm_final = 0;
	float rhs_norm = 0;

	bool success() const {
		return status == stop_reason::converged_atol ||
		       status == stop_reason::converged_rtol ||
		       status == stop_reason::converged_user;
	}
};

inline std::ostream & operator<<(std::ostream & os, solve_info::stop_reason r) {
	switch (r) {
	case solve_info::stop_reason::converged_atol: return os << "Converged to absolute tolerance";
	case solve_info::stop_reason::converged_rtol: return os << "Converged to residual tolerance";
	case solve_info::stop_reason::converged_user: return os << "Converged to user tolerance";
	case solve_info::stop_reason::diverged_dtol: return os << "Diverged: reached divergence tolerance";
	case solve_info::stop_reason::diverged_iters: return os << "Diverged: exceeded maximum iterations";
	case solve_info::stop_reason::diverged_breakdown: return os << "Diverged due to breakdown";
	default: return os << "Status unknown";
	}
}

}

#endif
