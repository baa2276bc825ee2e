// flecsolve/solvers/gmres.hh -- generic restarted GMRES.
//
// The reference's op::gmres (flecsolve/solvers/gmres.hh:38-361) operation
// for operation: left or right preconditioning (right by default), modified
// Gram-Schmidt against the stored basis, Givens rotations on a
// column-major (m+1)^2 Hessenberg matrix computed on the host (Bindel,
// Demmel, Kahan and Marques, "On computing Givens rotations reliably and
// efficiently", Algorithm 1), convergence on the rotated residual estimate
// |dw_{k+1}|, restart when k reaches max_krylov_dim, and the final
// back-substitution and correction.  The basis lives in the work array from
// index 3 on (work.data() + 3), as in the reference (gmres.hh:350-353).
#ifndef FLECSOLVE_AMD_SOLVERS_GMRES_HH
#define FLECSOLVE_AMD_SOLVERS_GMRES_HH

#include <algorithm>
#include <cmath>
#include <limits>
#include <memory>
#include <stdexcept>
#include <vector>

#include "flecsolve/solvers/krylov_parameters.hh"

namespace flecsolve::gmres {
inline constexpr int krylov_dim_bound = 100;
inline constexpr std::size_t nwork = (krylov_dim_bound + 1) + 3;
enum class precond_side { left, right };

struct settings : solver_settings {
	int max_krylov_dim = -1;
	precond_side pre_side = precond_side::right;
	bool restart = false;

	// reference gmres.hh:394-404 (flog_assert -> exception here)
	void validate() {
		if (max_krylov_dim < 0) max_krylov_dim = maxiter;
		if (max_krylov_dim > krylov_dim_bound)
			throw std::invalid_argument("GMRES: max_krylov_dim is larger than bound");
		if (!restart && maxiter > max_krylov_dim)
			throw std::invalid_argument("GMRES: maxiters must be less than or equal "
			                            "to max_krylov_dim when not using restart");
	}
};

// gmres.hh:368-380
inline std::istream & operator>>(std::istream & in, precond_side & s) {
	std::string tok;
	in >> tok;
	if (tok == "left") s = precond_side::left;
	else if (tok == "right") s = precond_side::right;
	else in.setstate(std::ios_base::failbit);
	return in;
}

// gmres.hh:406-421: the common keys plus max-krylov-dim (-1), pre-side
// (right) and restart (false)
struct options : solver_options {
	using settings_type = settings;
	options(const char * pre) : solver_options(pre) {}
	optdesc operator()(settings & s) {
		optdesc d = solver_options::operator()(s);
		d.with_default(label("max-krylov-dim"), &s.max_krylov_dim, -1)
			.with_default(label("pre-side"), &s.pre_side, precond_side::right)
			.with_default(label("restart"), &s.restart, false);
		return d;
	}
};
}

namespace flecsolve::op {

template<class Params>
struct gmres : base<Params, typename Params::input_var_t,
                    typename Params::output_var_t> {
	using base_t =
		base<Params, typename Params::input_var_t, typename Params::output_var_t>;
	using base_t::params;
	using real = typename Params::real;

	explicit gmres(Params p) : base_t(std::move(p)) { reset(); }
	const auto & get_operator() const { return params.A(); }

	void reset() {
		params.settings.validate();
		const auto & s = params.settings;
		dim = std::min(s.max_krylov_dim, s.maxiter);
		H.assign((std::size_t)(dim + 1) * (dim + 1), 0.0);
		cs.assign(dim + 1, 0.0);
		sn.assign(dim + 1, 0.0);
		dw.assign(dim + 1, 0.0);
		dy.assign(dim + 1, 0.0);
	}

	// On a device CSR with the identity or its Jacobi diagonal and device
	// vectors: the fused device driver (flecsolve/solvers/device_dispatch.hh);
	// otherwise the reference's sequence, generic().
	template<class DomainVec, class RangeVec>
	solve_info apply(const RangeVec & b, DomainVec & x) const {
		using drv = detail::fused_driver<detail::gmres_kind, Params, DomainVec, RangeVec>;
		if constexpr (drv::value) return drv::apply(params, b, x);
		else return generic(b, x);
	}

	template<class DomainVec, class RangeVec>
	solve_info generic(const RangeVec & b, DomainVec & x) const {
		using flecsolve::gmres::precond_side;
		const auto & A = params.A();
		const auto & P = params.P();
		auto & diag = params.ops.diagnostic;
		const auto & s = params.settings;
		auto & work = params.work;
		auto & res = work[0];
		auto & z = work[1];
		auto & v = work[2];
		auto * V = basis();
		solve_info info;

		real bnorm = b.l2norm().get();
		info.rhs_norm = bnorm;
		if (bnorm < std::numeric_limits<real>::epsilon()) bnorm = 1.0;
		const real tol = s.rtol * bnorm;
		const bool left = s.pre_side == precond_side::left;

		if (s.use_zero_guess) x.set_scalar(0.);
		auto initial_residual = [&](bool zero) {
			if (left) {
				if (zero) V[0].copy(b);
				else A.residual(b, x, V[0]);
				P.apply(V[0], res);
			}
			else {
				if (zero) res.copy(b);
				else A.residual(b, x, res);
			}
		};
		initial_residual(s.use_zero_guess);
		const real beta = res.l2norm().get();
		info.res_norm_initial = beta;
		if (beta < tol) {
			info.res_norm_final = beta;
			info.status = solve_info::stop_reason::converged_rtol;
			return info;
		}
		res.scale(1.0 / beta);
		V[0].copy(res);
		dw[0] = beta;
		real vnorm = beta;

		int k = 0;
		for (int it = 0; it < s.maxiter; ++it) {
			if (left) {
				A.apply(V[k], z);
				P.apply(z, v);
			}
			else {
				P.apply(V[k], z);
				A.apply(z, v);
			}
			// modified Gram-Schmidt (reference gmres.hh:265-279)
			for (int j = 0; j <= k; ++j) {
				const double h = v.dot(V[j]).get();
				v.axpy(-h, V[j], v);
				h_(j, k) = h;
			}
			h_(k + 1, k) = v.l2norm().get();

			vnorm = h_(k + 1, k);
			if (vnorm != 0.0) v.scale(1.0 / vnorm);
			V[k + 1].copy(v);
			for (int i = 0; i < k; ++i) rotate(i, k);
			if (vnorm != 0.0) {
				givens(k);
				rotate(k, k);
				h_(k + 1, k) = 0.0;
				const real t = dw[k];
				dw[k] = cs[k] * t;
				dw[k + 1] = -sn[k] * t;
			}
			vnorm = std::fabs(dw[k + 1]);
			++k;
			if (diag(x, vnorm)) {
				info.status = solve_info::stop_reason::converged_user;
				info.iters = it + 1;
				break;
			}
			if (vnorm < tol) {
				info.status = solve_info::stop_reason::converged_rtol;
				info.iters = it + 1;
				break;
			}
			if (k == s.max_krylov_dim && it != s.maxiter - 1) {
				back_solve(k - 1);
				correct(k - 1, P, z, v, x);
				initial_residual(false);
				const real br = res.l2norm().get();
				res.scale(1.0 / br);
				V[0].copy(res);
				dw[0] = br;
				++info.restarts;
				k = 0;
			}
		}
		if (k > 0) {
			back_solve(k - 1);
			correct(k - 1, P, z, v, x);
		}
		info.res_norm_final = vnorm;
		info.sol_norm_final = x.l2norm().get();
		if (info.iters == 0) info.status = solve_info::stop_reason::diverged_iters;
		return info;
	}

protected:
	real & h_(int i, int j) const { return H[(std::size_t)j * (dim + 1) + i]; }

	auto * basis() const {
		return params.work.data() +
		       (flecsolve::gmres::nwork - flecsolve::gmres::krylov_dim_bound - 1);
	}

	// reference gmres.hh:281-296
	void rotate(int i, int k) const {
		const real a = h_(i, k), b = h_(i + 1, k);
		h_(i, k) = cs[i] * a + sn[i] * b;
		h_(i + 1, k) = -sn[i] * a + cs[i] * b;
	}
	// reference gmres.hh:298-331
	void givens(int k) const {
		const real f = h_(k, k), g = h_(k + 1, k);
		real c, s;
		if (g == 0.0) {
			c = 1.0;
			s = 0.0;
		}
		else if (f == 0.0) {
			c = 0.0;
			s = (g < 0.0) ? -1.0 : 1.0;
		}
		else {
			real r = std::sqrt(f * f + g * g);
			r = 1.0 / r;
			c = std::fabs(f) * r;
			s = std::copysign(g * r, f);
		}
		cs[k] = c;
		sn[k] = s;
	}
	// reference gmres.hh:333-348
	void back_solve(int nr) const {
		dy[nr] = dw[nr] / h_(nr, nr);
		for (int k = nr - 1; k >= 0; --k) {
			dy[k] = dw[k];
			for (int i = k + 1; i <= nr; ++i) dy[k] -= h_(k, i) * dy[i];
			dy[k] = dy[k] / h_(k, k);
		}
	}
	// reference gmres.hh:244-263
	template<class Op, class W, class X>
	void correct(int nr, const Op & P, W & z, W & v, X & x) const {
		auto * V = basis();
		if (params.settings.pre_side == flecsolve::gmres::precond_side::right) {
			z.set_scalar(0.0);
			for (int i = 0; i <= nr; ++i) z.axpy(dy[i], V[i], z);
			P.apply(z, v);
			x.axpy(1.0, v, x);
		}
		else {
			for (int i = 0; i <= nr; ++i) x.axpy(dy[i], V[i], x);
		}
	}

	int dim = 0;
	mutable std::vector<real> H, cs, sn, dw, dy;
};

}

namespace flecsolve::gmres {
template<class Work>
struct solver : krylov_solver<op::gmres, settings, Work> {
	template<class W>
	solver(const settings & s, W && w)
		: krylov_solver<op::gmres, settings, Work>{s, std::forward<W>(w)} {}
};
template<class W>
solver(const settings &, W &&) -> solver<std::decay_t<W>>;
}

#endif
