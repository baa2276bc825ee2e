// flecsolve/vectors/core.hh -- vec::core<Data, Ops, Config>: the vector API
// the Krylov solvers are written against.
//
// Same template signature, member names and argument meaning as the
// reference's flecsolve/vectors/core.hh:26-353: every BLAS-1 method forwards
// to a static method of the Ops policy applied to the vector's data, and
// reductions return an object whose .get() yields the value.  Data and Ops are
// supplied by the storage backend (here vectors/device.hh: device-resident
// storage and HIP kernels behind the fsx C ABI).
#ifndef FLECSOLVE_AMD_VECTORS_CORE_HH
#define FLECSOLVE_AMD_VECTORS_CORE_HH

#include <cstddef>
#include <random>
#include <string_view>
#include <type_traits>
#include <utility>

#include "flecsolve/util/traits.hh"
#include "flecsolve/vectors/variable.hh"

namespace flecsolve::vec {

template<template<class> class Data, template<class> class Ops, class Config>
struct core;

namespace detail {
template<template<class> class D, template<class> class O, class C>
std::true_type is_core(const core<D, O, C> &);
std::false_type is_core(...);
}

template<class T>
inline constexpr bool is_vector_v =
	decltype(detail::is_core(std::declval<const std::decay_t<T> &>()))::value;

template<template<class> class Data, template<class> class Ops, class Config>
struct core : Config {
	using config = Config;
	using data_t = Data<Config>;
	using ops = Ops<data_t>;
	using scalar = typename Config::scalar;
	using real = typename Config::real;
	using len_t = typename Config::len_t;
	using var_t = typename Config::var_t;
	static constexpr auto var = Config::var;

	explicit core(data_t d) : data(std::move(d)) {}

	// ---- elementwise (reference core.hh:59-221) --------------------------
	template<class V, std::enable_if_t<is_vector_v<V>, bool> = true>
	void copy(const V & src) { ops::copy(src.data, data); }
	void zero() { ops::zero(data); }
	void set_scalar(scalar a) { ops::set_to_scalar(a, data); }
	template<class V, std::enable_if_t<is_vector_v<V>, bool> = true>
	void scale(scalar a, const V & x) { ops::scale(a, x.data, data); }
	void scale(scalar a) { ops::scale(a, data); }
	template<class X, class Y>
	void add(const X & x, const Y & y) { ops::add(x.data, y.data, data); }
	template<class X, class Y>
	void subtract(const X & x, const Y & y) { ops::subtract(x.data, y.data, data); }
	template<class X, class Y>
	void multiply(const X & x, const Y & y) { ops::multiply(x.data, y.data, data); }
	template<class X, class Y>
	void divide(const X & x, const Y & y) { ops::divide(x.data, y.data, data); }
	template<class X>
	void reciprocal(const X & x) { ops::reciprocal(x.data, data); }
	template<class X, class Y>
	void linear_sum(scalar a, const X & x, scalar b, const Y & y) {
		ops::linear_sum(a, x.data, b, y.data, data);
	}
	// this = alpha * x + y
	template<class X, class Y>
	void axpy(scalar a, const X & x, const Y & y) {
		ops::axpy(a, x.data, y.data, data);
	}
	// this = alpha * x + beta * this
	template<class X>
	void axpby(scalar a, scalar b, const X & x) { ops::axpby(a, b, x.data, data); }
	template<class X>
	void abs(const X & x) { ops::abs(x.data, data); }
	template<class X>
	void add_scalar(const X & x, scalar a) { ops::add_scalar(x.data, a, data); }

	// ---- reductions (reference core.hh:229-290) ---------------------------
	auto min() const { return ops::min(data); }
	auto max() const { return ops::max(data); }
	auto l1norm() const { return ops::template lp_norm<1>(data); }
	auto l2norm() const { return ops::template lp_norm<2>(data); }
	template<unsigned short p>
	auto lp_norm() const { return ops::template lp_norm<p>(data); }
	auto inf_norm() const { return ops::inf_norm(data); }
	template<class X>
	auto dot(const X & x) const { return ops::dot(data, x.data); }
	auto global_size() const { return ops::global_size(data); }
	std::size_t local_size() const { return ops::local_size(data); }

	// element access (reference core.hh:338-340): what Ops::retreive returns,
	// a reference into host storage or, for device storage, a proxy that
	// reads or writes the one element
	decltype(auto) operator[](len_t i) { return ops::retreive(data, i); }
	decltype(auto) operator[](len_t i) const { return ops::retreive(data, i); }

	void set_random() { ops::set_random(data, std::random_device{}()); }
	void set_random(unsigned seed) { ops::set_random(data, seed); }
	void dump(std::string_view name) { ops::dump(name, data); }

	// single-field vectors: subsetting by the own variable is the identity
	template<auto ov>
	constexpr decltype(auto) subset(variable_t<ov>) const { return *this; }
	template<auto ov>
	constexpr decltype(auto) subset(variable_t<ov>) { return *this; }

	data_t data;
};

template<template<class> class D, template<class> class O, class C>
bool operator==(const core<D, O, C> & a, const core<D, O, C> & b) {
	return a.data == b.data;
}
template<template<class> class D, template<class> class O, class C>
bool operator!=(const core<D, O, C> & a, const core<D, O, C> & b) {
	return !(a.data == b.data);
}

}

namespace flecsolve {
using vec::is_vector_v;
}

#endif
