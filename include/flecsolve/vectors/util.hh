// flecsolve/vectors/util.hh -- vector helpers: norm selection and work
// arrays shaped like an exemplar.
//
// norm_type and its stream extraction follow the reference's
// flecsolve/vectors/util.hh:38-40 and util.cc:8-27 (inf, l2, l1, any case).
// make_work<N>(v) builds N vectors shaped like v (the reference's
// work_factory, solvers/solver_settings.hh:319-330): the Data policy's
// make_like() allocates a new buffer of the same length on the same context.
#ifndef FLECSOLVE_AMD_VECTORS_UTIL_HH
#define FLECSOLVE_AMD_VECTORS_UTIL_HH

#include <algorithm>
#include <array>
#include <cctype>
#include <cstddef>
#include <istream>
#include <string>
#include <utility>

namespace flecsolve::vec {

enum class norm_type { inf, l2, l1 };

inline std::istream & operator>>(std::istream & in, norm_type & n) {
	std::string tok;
	in >> tok;
	std::transform(tok.begin(), tok.end(), tok.begin(),
	               [](unsigned char c) { return (char)std::tolower(c); });
	if (tok == "inf") n = norm_type::inf;
	else if (tok == "l1") n = norm_type::l1;
	else if (tok == "l2") n = norm_type::l2;
	else in.setstate(std::ios_base::failbit);
	return in;
}

namespace detail {
template<class V, std::size_t... I>
std::array<V, sizeof...(I)> make_work(const V & v, std::index_sequence<I...>) {
	return {((void)I, V(v.data.make_like()))...};
}
}

template<std::size_t N, class V>
std::array<V, N> make_work(const V & v) {
	return detail::make_work(v, std::make_index_sequence<N>{});
}

}

#endif
