// flecsolve/solvers/krylov_parameters.hh -- binding of operator, preconditioner,
// diagnostic, settings and work space into a solver operator.
//
// Interface of the reference's flecsolve/solvers/krylov_parameters.hh:29-126:
//   solver(settings, work)(op::ref(A) [, P [, diagnostic]]) -> op::core<...>
// The preconditioner defaults to the identity (krylov_parameters.hh:65-70),
// the diagnostic to "never stop" (29-31); binding an rvalue solver moves the
// settings and work, binding an lvalue copies them (99-112).  The work array
// must be indexable with .data() and tuple-like (std::array of vectors).
#ifndef FLECSOLVE_AMD_SOLVERS_KRYLOV_PARAMETERS_HH
#define FLECSOLVE_AMD_SOLVERS_KRYLOV_PARAMETERS_HH

#include <type_traits>
#include <utility>

#include "flecsolve/operators/handle.hh"
#include "flecsolve/operators/shell.hh"
#include "flecsolve/solvers/solver_settings.hh"
#include "flecsolve/util/traits.hh"

namespace flecsolve {

template<class Op, class Precond, class Diag>
struct krylov_ops {
	op::handle<Op> A;
	op::handle<Precond> P;
	Diag diagnostic;
};

template<class Settings, class Work, class Op, class Precond, class Diag>
struct krylov_parameters {
	using work_t = Work;
	using workvec_t = typename std::remove_reference_t<Work>::value_type;
	using real = typename workvec_t::real;
	using scalar = typename workvec_t::scalar;
	using input_var_t = std::decay_t<decltype(Op::input_var)>;
	using output_var_t = std::decay_t<decltype(Op::output_var)>;

	template<class S, class W, class D>
	krylov_parameters(S && s, W && w, op::handle<Op> a, op::handle<Precond> p,
	                  D && d)
		: work(std::forward<W>(w)), settings(std::forward<S>(s)),
		  ops{a, p, std::forward<D>(d)} {}

	const Op & A() const { return ops.A.get(); }
	const Precond & P() const { return ops.P.get(); }

	mutable Work work;
	Settings settings;
	mutable krylov_ops<Op, Precond, Diag> ops;
};

namespace detail {
using identity_handle = std::decay_t<decltype(op::I)>;
using default_diag_t = std::decay_t<decltype(default_diagnostic)>;

template<class S, class W, class A>
auto make_params(S && s, W && w, op::handle<A> a) {
	return krylov_parameters<std::decay_t<S>, std::decay_t<W>, A,
	                         typename identity_handle::type, default_diag_t>(
		std::forward<S>(s), std::forward<W>(w), a, op::I, default_diagnostic);
}
template<class S, class W, class A, class P>
auto make_params(S && s, W && w, op::handle<A> a, op::handle<P> p) {
	return krylov_parameters<std::decay_t<S>, std::decay_t<W>, A, P,
	                         default_diag_t>(std::forward<S>(s),
	                                         std::forward<W>(w), a, p,
	                                         default_diagnostic);
}
template<class S, class W, class A, class P, class D>
auto make_params(S && s, W && w, op::handle<A> a, op::handle<P> p, D && d) {
	return krylov_parameters<std::decay_t<S>, std::decay_t<W>, A, P,
	                         std::decay_t<D>>(std::forward<S>(s),
	                                          std::forward<W>(w), a, p,
	                                          std::forward<D>(d));
}
}

namespace detail {
// A fused device driver a bound solver's apply() may hand the solve to.
// Specialised by flecsolve/solvers/device_dispatch.hh (included with the
// device CSR operator) for a device CSR, the identity or Jacobi
// preconditioner and device vectors; value == false keeps the generic
// sequence.
struct cg_kind {};
struct bicgstab_kind {};
struct gmres_kind {};
template<class Kind, class Params, class DomainVec, class RangeVec, class = void>
struct fused_driver : std::false_type {};
}

template<template<class> class OpType, class Settings, class Workspace>
struct krylov_solver {
	using settings_type = Settings;

	template<class... Ops>
	auto bind(Ops &&... ops) && {
		auto prm = detail::make_params(std::move(settings), std::move(workspace),
		                               std::forward<Ops>(ops)...);
		return OpType<decltype(prm)>(std::move(prm));
	}
	template<class... Ops>
	auto bind(Ops &&... ops) & {
		auto prm = detail::make_params(settings, workspace,
		                               std::forward<Ops>(ops)...);
		return OpType<decltype(prm)>(std::move(prm));
	}
	template<class... Ops>
	auto operator()(Ops &&... ops) && {
		return op::make(std::move(*this).bind(std::forward<Ops>(ops)...));
	}
	template<class... Ops>
	auto operator()(Ops &&... ops) & {
		return op::make(bind(std::forward<Ops>(ops)...));
	}

	settings_type settings;
	Workspace workspace;
};

}

#endif
