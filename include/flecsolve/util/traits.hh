// flecsolve/util/traits.hh -- numeric and wrapper traits used by the vector,
// operator and solver layers.  Same names and meaning as the reference's
// flecsolve/util/traits.hh:29-62 (num_traits, is_reference_wrapper).
#ifndef FLECSOLVE_AMD_UTIL_TRAITS_HH
#define FLECSOLVE_AMD_UTIL_TRAITS_HH

#include <complex>
#include <functional>
#include <memory>
#include <type_traits>

namespace flecsolve {

template<class T>
struct num_traits {
	using scalar = T;
	using real = T;
	static constexpr bool is_complex = false;
};
template<class T>
struct num_traits<std::complex<T>> {
	using scalar = std::complex<T>;
	using real = T;
	static constexpr bool is_complex = true;
};

template<class T>
struct is_reference_wrapper : std::false_type {};
template<class T>
struct is_reference_wrapper<std::reference_wrapper<T>> : std::true_type {};
template<class T>
inline constexpr bool is_reference_wrapper_v = is_reference_wrapper<T>::value;

// reference solvers/krylov_parameters.hh:29-31: the default diagnostic never
// stops the iteration
inline constexpr auto default_diagnostic = [](auto &&, double) { return false; };

}

#endif
