// Host-memory Data / Ops policies for vec::core and a host CSR operator.
//
// Test infrastructure: lets the C++ header API (time integrators, generic
// Krylov solvers, factory) run on the CPU, so that its logic is checked
// against the Python oracle without a GPU.  Each operation is the same
// elementwise formula as the device kernel of the same name (a*x + y,
// a*x + b*y, ..., two roundings, no contraction), and every reduction sums in
// index order.
#ifndef FLECSOLVE_AMD_TESTS_HOST_VECTOR_HH
#define FLECSOLVE_AMD_TESTS_HOST_VECTOR_HH

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <string_view>
#include <vector>

#include "flecsolve/operators/core.hh"
#include "flecsolve/vectors/core.hh"

namespace flecsolve::vec {

template<class T>
struct host_config {
	using scalar = T;
	using real = T;
	using len_t = std::size_t;
	using var_t = anon_var;
	static constexpr auto var = variable<anon_var::anonymous>;
	static constexpr std::size_t num_components = 1;
	static constexpr bool is_complex = false;
};

template<class Config>
struct host_data {
	using config = Config;
	using scalar = typename Config::scalar;
	host_data() = default;
	explicit host_data(std::size_t n) : s(std::make_shared<std::vector<scalar>>(n, scalar(0))) {}
	host_data make_like() const { return host_data(s->size()); }
	std::size_t size() const { return s->size(); }
	scalar * p() const { return s->data(); }
	friend bool operator==(const host_data & a, const host_data & b) { return a.s == b.s; }
	friend bool operator!=(const host_data & a, const host_data & b) { return a.s != b.s; }
	std::shared_ptr<std::vector<scalar>> s;
};

template<class T>
struct host_ready {
	T val;
	T get() const { return val; }
};

template<class Data>
struct host_ops {
	using scalar = typename Data::scalar;
	using real = scalar;
	using len_t = std::size_t;
	template<class F>
	static void each(const Data & d, F f) {
		for (std::size_t i = 0; i < d.size(); ++i) f(i);
	}
	static void copy(const Data & s, Data & d) {
		auto * a = s.p(); auto * b = d.p();
		each(d, [&](std::size_t i) { b[i] = a[i]; });
	}
	static void zero(Data & x) { set_to_scalar(0, x); }
	static void set_to_scalar(scalar a, Data & x) {
		auto * p = x.p();
		each(x, [&](std::size_t i) { p[i] = a; });
	}
	static void scale(scalar a, Data & x) {
		auto * p = x.p();
		each(x, [&](std::size_t i) { p[i] = a * p[i]; });
	}
	static void scale(scalar a, const Data & x, Data & y) {
		auto * s = x.p(); auto * d = y.p();
		each(y, [&](std::size_t i) { d[i] = a * s[i]; });
	}
	static void add(const Data & x, const Data & y, Data & z) {
		auto * a = x.p(); auto * b = y.p(); auto * c = z.p();
		each(z, [&](std::size_t i) { c[i] = a[i] + b[i]; });
	}
	static void subtract(const Data & x, const Data & y, Data & z) {
		auto * a = x.p(); auto * b = y.p(); auto * c = z.p();
		each(z, [&](std::size_t i) { c[i] = a[i] - b[i]; });
	}
	static void multiply(const Data & x, const Data & y, Data & z) {
		auto * a = x.p(); auto * b = y.p(); auto * c = z.p();
		each(z, [&](std::size_t i) { c[i] = a[i] * b[i]; });
	}
	static void divide(const Data & x, const Data & y, Data & z) {
		auto * a = x.p(); auto * b = y.p(); auto * c = z.p();
		each(z, [&](std::size_t i) { c[i] = a[i] / b[i]; });
	}
	static void reciprocal(const Data & x, Data & y) {
		auto * a = x.p(); auto * b = y.p();
		each(y, [&](std::size_t i) { b[i] = scalar(1) / a[i]; });
	}
	static void linear_sum(scalar a, const Data & x, scalar b, const Data & y, Data & z) {
		auto * u = x.p(); auto * v = y.p(); auto * w = z.p();
		each(z, [&](std::size_t i) { w[i] = a * u[i] + b * v[i]; });
	}
	static void axpy(scalar a, const Data & x, const Data & y, Data & z) {
		auto * u = x.p(); auto * v = y.p(); auto * w = z.p();
		each(z, [&](std::size_t i) { w[i] = a * u[i] + v[i]; });
	}
	static void axpby(scalar a, scalar b, const Data & x, Data & z) {
		auto * u = x.p(); auto * w = z.p();
		each(z, [&](std::size_t i) { w[i] = a * u[i] + b * w[i]; });
	}
	static void abs(const Data & x, Data & y) {
		auto * a = x.p(); auto * b = y.p();
		each(y, [&](std::size_t i) { b[i] = std::fabs(a[i]); });
	}
	static void add_scalar(const Data & x, scalar a, Data & y) {
		auto * s = x.p(); auto * d = y.p();
		each(y, [&](std::size_t i) { d[i] = s[i] + a; });
	}
	static void set_random(Data &, unsigned) {}
	static void dump(std::string_view, const Data &) {}
	static host_ready<real> min(const Data & x) {
		return {*std::min_element(x.p(), x.p() + x.size())};
	}
	static host_ready<real> max(const Data & x) {
		return {*std::max_element(x.p(), x.p() + x.size())};
	}
	template<unsigned short p>
	static host_ready<real> lp_norm(const Data & x) {
		real s = 0;
		for (std::size_t i = 0; i < x.size(); ++i)
			s += (p == 1) ? std::fabs(x.p()[i]) : x.p()[i] * x.p()[i];
		return {p == 1 ? s : std::sqrt(s)};
	}
	static host_ready<real> inf_norm(const Data & x) {
		real m = 0;
		for (std::size_t i = 0; i < x.size(); ++i) m = std::max(m, std::fabs(x.p()[i]));
		return {m};
	}
	static host_ready<scalar> dot(const Data & x, const Data & y) {
		scalar s = 0;
		for (std::size_t i = 0; i < x.size(); ++i) s += x.p()[i] * y.p()[i];
		return {s};
	}
	static host_ready<len_t> global_size(const Data & x) { return {x.size()}; }
	static len_t local_size(const Data & x) { return x.size(); }
};

template<class T = double>
using host = core<host_data, host_ops, host_config<T>>;

template<class T = double>
host<T> make_host(std::size_t n) {
	return host<T>(host_data<host_config<T>>(n));
}

}

namespace flecsolve::mat {

// y = A x, each row summed in CSR order from 0
struct host_csr : op::base<> {
	host_csr(std::vector<int32_t> rp, std::vector<int32_t> ci, std::vector<double> v)
		: rp(std::move(rp)), ci(std::move(ci)), v(std::move(v)) {}
	template<class D, class R>
	void apply(const D & x, R & y) const {
		const double * xp = x.data.p();
		double * yp = y.data.p();
		for (std::size_t r = 0; r + 1 < rp.size(); ++r) {
			double s = 0;
			for (int32_t e = rp[r]; e < rp[r + 1]; ++e) s += v[e] * xp[ci[e]];
			yp[r] = s;
		}
	}
	std::vector<int32_t> rp, ci;
	std::vector<double> v;
};

}

#endif
