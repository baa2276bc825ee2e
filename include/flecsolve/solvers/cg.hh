// flecsolve/solvers/cg.hh -- generic preconditioned CG and flexible CG.
//
// The iteration is the reference's op::cg::apply
// (flecsolve/solvers/cg.hh:43-139) operation for operation, over any vector
// type with the vec::core interface: the same tolerance (float rtol times
// |b|, |b| == 0 -> 1), the same order of the diagnostic and tolerance tests,
// the recursive residual, and iters == 0 with diverged_iters when maxiter is
// exhausted.  op::fcg follows the reference's flexible variant
// (cg.hh:161-243).  Over vec::device each reduction's .get() is one device
// reduction read back to the host, the reference's blocking point; the fused
// graph-captured driver with device-side scalars is fsx_cg_solve.
#ifndef FLECSOLVE_AMD_SOLVERS_CG_HH
#define FLECSOLVE_AMD_SOLVERS_CG_HH

#include <array>
#include <cstddef>

#include "flecsolve/solvers/krylov_parameters.hh"

namespace flecsolve::op {

template<class Params>
struct cg : base<Params, typename Params::input_var_t,
                 typename Params::output_var_t> {
	using base_t =
		base<Params, typename Params::input_var_t, typename Params::output_var_t>;
	using base_t::params;
	explicit cg(Params p) : base_t(std::move(p)) {}
	const auto & get_operator() const { return params.A(); }
	long negative_curvature = 0; // the reference logs these (cg.hh:101-104)

	// On a device CSR with the identity or its Jacobi diagonal and device
	// vectors: the fused device driver (flecsolve/solvers/device_dispatch.hh);
	// otherwise the reference's sequence, generic().
	template<class DomainVec, class RangeVec>
	solve_info apply(const RangeVec & b, DomainVec & x) const {
		using drv = detail::fused_driver<detail::cg_kind, Params, DomainVec, RangeVec>;
		if constexpr (drv::value) return drv::apply(params, b, x);
		else return generic(b, x);
	}

	template<class DomainVec, class RangeVec>
	solve_info generic(const RangeVec & b, DomainVec & x) const {
		using scalar = typename DomainVec::scalar;
		using real = typename DomainVec::real;
		solve_info info;
		const auto & A = params.A();
		const auto & P = params.P();
		auto & diag = params.ops.diagnostic;
		const auto & s = params.settings;
		auto & [r, z, p, w] = params.work;

		real bnorm = b.l2norm().get();
		if (bnorm == 0.0) bnorm = 1.0;
		const real tol = s.rtol * bnorm;
		info.rhs_norm = bnorm;

		if (s.use_zero_guess) {
			info.sol_norm_initial = 0;
			x.set_scalar(0.);
			r.copy(b);
		}
		else {
			info.sol_norm_initial = x.l2norm().get();
			A.residual(b, x, r);
		}
		real res = r.l2norm().get();
		if (res < tol) {
			info.res_norm_initial = res;
			info.res_norm_final = res;
			info.status = solve_info::stop_reason::converged_rtol;
			return info;
		}

		P.apply(r, z);
		scalar rho_old = 2.0, rho = z.dot(r).get();
		rho_old = rho;
		p.copy(z);
		for (int it = 0; it < s.maxiter; ++it) {
			A.apply(p, w);
			scalar alpha = w.dot(p).get();
			if (alpha <= 0.0) ++const_cast<cg *>(this)->negative_curvature;
			alpha = rho / alpha;
			x.axpy(alpha, p, x);
			r.axpy(-alpha, w, r);
			res = r.l2norm().get();
			if (diag(x, res)) {
				info.iters = it + 1;
				info.status = solve_info::stop_reason::converged_user;
				break;
			}
			if (res < tol) {
				info.iters = it + 1;
				info.status = solve_info::stop_reason::converged_rtol;
				break;
			}
			P.apply(r, z);
			rho_old = rho;
			rho = r.dot(z).get();
			const scalar beta = rho / rho_old;
			p.axpy(beta, p, z);
		}
		info.res_norm_final = res;
		info.sol_norm_final = x.l2norm().get();
		if (info.iters == 0) info.status = solve_info::stop_reason::diverged_iters;
		return info;
	}
};

template<class Params>
struct fcg : base<Params, typename Params::input_var_t,
                  typename Params::output_var_t> {
	using base_t =
		base<Params, typename Params::input_var_t, typename Params::output_var_t>;
	using base_t::params;
	explicit fcg(Params p) : base_t(std::move(p)) {}
	const auto & get_operator() const { return params.A(); }

	template<class DomainVec, class RangeVec>
	solve_info apply(const RangeVec & b, DomainVec & u) const {
		using scalar = typename DomainVec::scalar;
		using real = typename DomainVec::real;
		solve_info info;
		const auto & A = params.A();
		const auto & P = params.P();
		auto & diag = params.ops.diagnostic;
		const auto & s = params.settings;
		auto & [r, v, w, q, d] = params.work;

		real bnorm = b.l2norm().get();
		if (bnorm == 0.0) bnorm = 1.0;
		const real tol = s.rtol * bnorm;
		info.rhs_norm = bnorm;
		if (s.use_zero_guess) {
			info.sol_norm_initial = 0;
			u.set_scalar(0.);
			r.copy(b);
		}
		else {
			info.sol_norm_initial = u.l2norm().get();
			A.residual(b, u, r);
		}
		real res = r.l2norm().get();
		if (res < tol) {
			info.res_norm_initial = res;
			info.res_norm_final = res;
			info.status = solve_info::stop_reason::converged_rtol;
			return info;
		}
		scalar rho = 0;
		for (int it = 0; it < s.maxiter; ++it) {
			P.apply(r, v);
			A.apply(v, w);
			const scalar alpha = v.dot(r).get();
			const scalar beta = v.dot(w).get();
			if (it > 0) {
				const scalar gamma = v.dot(q).get();
				d.axpy((-gamma) / rho, d, v);
				q.axpy((-gamma) / rho, q, w);
				rho = beta - (gamma * gamma) / rho;
			}
			else {
				d.copy(v);
				q.copy(w);
				rho = beta;
			}
			u.axpy(alpha / rho, d, u);
			r.axpy(-alpha / rho, q, r);
			res = r.l2norm().get();
			if (diag(u, res)) {
				info.iters = it + 1;
				info.status = solve_info::stop_reason::converged_user;
				break;
			}
			if (res < tol) {
				info.iters = it + 1;
				info.status = solve_info::stop_reason::converged_rtol;
				break;
			}
		}
		info.res_norm_final = res;
		info.sol_norm_final = u.l2norm().get();
		if (info.iters == 0) info.status = solve_info::stop_reason::diverged_iters;
		return info;
	}
};

}

namespace flecsolve::cg {
inline constexpr std::size_t nwork = 4; // r, z, p, w
using settings = solver_settings;
using options = solver_options;

template<class Work>
struct solver : krylov_solver<op::cg, settings, Work> {
	template<class W>
	solver(const settings & s, W && w)
		: krylov_solver<op::cg, settings, Work>{s, std::forward<W>(w)} {}
};
template<class W>
solver(const settings &, W &&) -> solver<std::decay_t<W>>;
}

namespace flecsolve::fcg {
inline constexpr std::size_t nwork = 5; // r, v, w, q, d
using settings = solver_settings;

template<class Work>
struct solver : krylov_solver<op::fcg, settings, Work> {
	template<class W>
	solver(const settings & s, W && w)
		: krylov_solver<op::fcg, settings, Work>{s, std::forward<W>(w)} {}
};
template<class W>
solver(const settings &, W &&) -> solver<std::decay_t<W>>;
}

#endif
