// flecsolve/operators/shell.hh -- operators from callables, and the identity.
// Interface of the reference's flecsolve/operators/shell.hh:24-95 (shell,
// make_shell, make_shared_shell, make_identity, I): the identity
// preconditioner copies its input (y.copy(x)).
#ifndef FLECSOLVE_AMD_OPERATORS_SHELL_HH
#define FLECSOLVE_AMD_OPERATORS_SHELL_HH

#include <utility>

#include "flecsolve/operators/core.hh"
#include "flecsolve/operators/handle.hh"

namespace flecsolve::op {

template<class F,
         class ivar = variable_t<anon_var::anonymous>,
         class ovar = variable_t<anon_var::anonymous>>
struct shell : base<std::nullptr_t, ivar, ovar> {
	explicit shell(F f) : f(std::move(f)) {}
	template<class D, class R>
	decltype(auto) apply(const D & x, R & y) const { return f(x, y); }

protected:
	F f;
};

template<class F>
auto make_shell(F && f) {
	return core<shell<std::decay_t<F>>>(std::forward<F>(f));
}
template<class F>
auto make_shared_shell(F && f) {
	return make_shared<shell<std::decay_t<F>>>(std::forward<F>(f));
}

inline auto make_identity() {
	return make_shared_shell([](const auto & x, auto & y) { y.copy(x); });
}
template<auto iv, auto ov>
auto make_identity(variable_t<iv>, variable_t<ov>) { return make_identity(); }

inline const auto I = make_identity();

}

#endif
