// flecsolve/solvers/bicgstab.hh -- generic preconditioned BiCGSTAB.
//
// The reference's op::bicgstab::apply (flecsolve/solvers/bicgstab.hh:39-187)
// operation for operation, including its quirks: rho starts as {2.0, 1.0}
// (list-initialised std::vector, bicgstab.hh:89); on breakdown
// (sqrt|rho| < eps |r~|) the residual is recomputed and r~, p reset before
// `continue` (100-113); the early exit compares |s| with the absolute rtol and
// reports iters = iter (134-143); omega == 0 stops with diverged_breakdown.
// A zero r~.v, which the reference asserts against (129), throws here.
#ifndef FLECSOLVE_AMD_SOLVERS_BICGSTAB_HH
#define FLECSOLVE_AMD_SOLVERS_BICGSTAB_HH

#include <cmath>
#include <limits>
#include <stdexcept>

#include "flecsolve/solvers/krylov_parameters.hh"

namespace flecsolve::op {

template<class Params>
struct bicgstab : base<Params, typename Params::input_var_t,
                       typename Params::output_var_t> {
	using base_t =
		base<Params, typename Params::input_var_t, typename Params::output_var_t>;
	using base_t::params;
	using real = typename Params::real;
	explicit bicgstab(Params p) : base_t(std::move(p)) {}
	const auto & get_operator() const { return params.A(); }

	// On a device CSR with the identity or its Jacobi diagonal and device
	// vectors: the fused device driver (flecsolve/solvers/device_dispatch.hh);
	// otherwise the reference's sequence, generic().
	template<class DomainVec, class RangeVec>
	solve_info apply(const RangeVec & b, DomainVec & x) const {
		using drv = detail::fused_driver<detail::bicgstab_kind, Params, DomainVec, RangeVec>;
		if constexpr (drv::value) return drv::apply(params, b, x);
		else return generic(b, x);
	}

	template<class DomainVec, class RangeVec>
	solve_info generic(const RangeVec & b, DomainVec & x) const {
		solve_info info;
		const auto & A = params.A();
		const auto & P = params.P();
		auto & diag = params.ops.diagnostic;
		const auto & set = params.settings;
		auto & [res, rt, p, v, ph, s, sh, t] = params.work;

		real bnorm = b.l2norm().get();
		if (bnorm == 0.) bnorm = 1.;
		const real tol = set.rtol * bnorm;
		info.rhs_norm = bnorm;
		if (set.use_zero_guess) {
			info.sol_norm_initial = 0;
			res.copy(b);
			x.set_scalar(0.);
		}
		else {
			info.sol_norm_initial = x.l2norm().get();
			A.residual(b, x, res);
		}
		real rnorm = res.l2norm().get();
		real rt_norm = rnorm;
		info.res_norm_initial = rnorm;
		if (rnorm < tol) {
			info.status = solve_info::stop_reason::converged_rtol;
			info.res_norm_final = rnorm;
			return info;
		}

		real alpha = 1.0, beta = 0.0, omega = 1.0;
		real rho_prev = 2.0, rho = 1.0;
		rt.copy(res);
		p.zero();
		v.zero();
		for (int it = 0; it < set.maxiter; ++it) {
			rho = rt.dot(res).get();
			if (std::sqrt(std::fabs(rho)) <
			    std::numeric_limits<real>::epsilon() * rt_norm) {
				// breakdown: restart from the true residual
				A.residual(b, x, res);
				rt.copy(res);
				rnorm = res.l2norm().get();
				rho = rt_norm = rnorm;
				p.copy(res);
				++info.restarts;
				continue;
			}
			if (it == 0)
				p.copy(res);
			else {
				beta = (rho / rho_prev) * (alpha / omega);
				p.axpy(-omega, v, p);
				p.axpy(beta, p, res);
			}
			P.apply(p, ph);
			A.apply(ph, v);
			alpha = rt.dot(v).get();
			if (alpha == 0.) throw std::runtime_error("BiCGSTAB: encountered alpha = 0");
			alpha = rho / alpha;
			s.axpy(-alpha, v, res);
			const real snorm = s.l2norm().get();
			if (snorm < set.rtol) {
				x.axpy(alpha, ph, x);
				info.iters = it;
				info.status = solve_info::stop_reason::converged_rtol;
				break;
			}
			P.apply(s, sh);
			A.apply(sh, t);
			const real tt = t.dot(t).get();
			const real ts = t.dot(s).get();
			omega = (tt == 0.0) ? 0.0 : ts / tt;
			x.axpy(alpha, ph, x);
			x.axpy(omega, sh, x);
			res.axpy(-omega, t, s);
			rnorm = res.l2norm().get();
			if (diag(x, rnorm)) {
				info.status = solve_info::stop_reason::converged_user;
				info.iters = it + 1;
				break;
			}
			if (rnorm < tol) {
				info.status = solve_info::stop_reason::converged_rtol;
				info.iters = it + 1;
				break;
			}
			if (omega == 0.0) {
				info.iters = it + 1;
				info.status = solve_info::stop_reason::diverged_breakdown;
				break;
			}
			rho_prev = rho;
		}
		info.res_norm_final = rnorm;
		info.sol_norm_final = x.l2norm().get();
		if (info.iters == 0) info.status = solve_info::stop_reason::diverged_iters;
		return info;
	}
};

}

namespace flecsolve::bicgstab {
inline constexpr std::size_t nwork = 8; // res, r~, p, v, p^, s, s^, t
struct settings : solver_settings {};
struct options : solver_options {
	using settings_type = settings;
	options(const char * pre) : solver_options(pre) {}
	optdesc operator()(settings & s) { return solver_options::operator()(s); }
};

template<class Work>
struct solver : krylov_solver<op::bicgstab, settings, Work> {
	template<class W>
	solver(const settings & s, W && w)
		: krylov_solver<op::bicgstab, settings, Work>{s, std::forward<W>(w)} {}
};
template<class W>
solver(const settings &, W &&) -> solver<std::decay_t<W>>;
}

#endif
