// flecsolve/vectors/variable.hh -- variable tags that name the physics field a
// vector or operator acts on.  Interface of the reference's
// flecsolve/vectors/variable.hh:24-62 (anon_var, variable_t, variable,
// multivariable_t); single-field solves use anon_var::anonymous throughout.
#ifndef FLECSOLVE_AMD_VECTORS_VARIABLE_HH
#define FLECSOLVE_AMD_VECTORS_VARIABLE_HH

#include <cstddef>
#include <limits>

namespace flecsolve {

enum class anon_var : std::size_t {
	anonymous = std::numeric_limits<std::size_t>::max()
};

template<auto V>
struct variable_t {
	static constexpr auto value = V;
};
template<auto A, auto B>
constexpr bool operator==(variable_t<A>, variable_t<B>) { return A == B; }
template<auto A, auto B>
constexpr bool operator!=(variable_t<A>, variable_t<B>) { return !(A == B); }

template<auto V>
inline constexpr variable_t<V> variable{};

template<auto... Vs>
struct multivariable_t {};
template<auto... Vs>
inline constexpr multivariable_t<Vs...> multivariable{};

}

#endif
