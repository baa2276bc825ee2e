// flecsolve/operators/core.hh -- operator policy base and op::core.
//
// Interface of the reference's flecsolve/operators/core.hh:29-148: op::base
// holds parameters and the input/output variables; op::core<P> adds apply(),
// operator() and residual() (r = b - A x computed as apply then in-place
// subtract, core.hh:135-142); op::make wraps a policy.
#ifndef FLECSOLVE_AMD_OPERATORS_CORE_HH
#define FLECSOLVE_AMD_OPERATORS_CORE_HH

#include <cstddef>
#include <type_traits>
#include <utility>

#include "flecsolve/vectors/core.hh"
#include "flecsolve/vectors/variable.hh"

namespace flecsolve::op {

template<class Params = std::nullptr_t,
         class ivar = variable_t<anon_var::anonymous>,
         class ovar = variable_t<anon_var::anonymous>>
struct base {
	static constexpr auto input_var = ivar{};
	static constexpr auto output_var = ovar{};
	using params_t = Params;

	base() = default;
	template<class... A,
	         std::enable_if_t<std::is_constructible_v<Params, A &&...>, bool> = true>
	explicit base(A &&... a) : params(std::forward<A>(a)...) {}

	params_t & get_params() { return params; }
	const params_t & get_params() const { return params; }

protected:
	params_t params{};
};

template<class P>
struct core : P {
	static constexpr auto input_var = P::input_var;
	static constexpr auto output_var = P::output_var;
	using policy_type = P;

	template<class... A,
	         std::enable_if_t<std::is_constructible_v<P, A &&...>, bool> = true>
	core(A &&... a) : P(std::forward<A>(a)...) {}

	template<class D, class R>
	decltype(auto) apply(const D & x, R & y) const {
		return P::apply(x.subset(input_var), y.subset(output_var));
	}
	template<class D, class R>
	decltype(auto) operator()(const D & x, R & y) const {
		return apply(x, y);
	}
	template<class B, class X, class R>
	void residual(const B & b, const X & x, R & r) const {
		apply(x, r);
		auto & rs = r.subset(output_var);
		rs.subtract(b.subset(output_var), rs);
	}
};

template<class P>
auto make(P && p) {
	return core<std::decay_t<P>>(std::forward<P>(p));
}

template<class T>
struct is_operator : std::false_type {};
template<class P>
struct is_operator<core<P>> : std::true_type {};
template<class T>
inline constexpr bool is_operator_v = is_operator<T>::value;

}

#endif
