// flecsolve/vectors/device.hh -- device-resident vector: Data and Ops
// policies of vec::core backed by libfsx (include/fsx.h).
//
// Replaces the reference's FleCSI-field vector (vectors/topo_view.hh,
// vectors/data/topo_view.hh, vectors/operations/topo_view.hh): Data is a
// shallow, reference-counted handle to a device buffer (copy = alias,
// == compares buffer identity, as topo_view compares field ids,
// data/topo_view.hh:73-81); each Ops method is one HIP kernel launched on the
// context stream; reductions return a future-like object whose .get() is the
// host value (the reference's .get() also blocks, topo_view.hh:265-269).
//
// Storage is allocated on first use, so a work array sized for the largest
// Krylov space (GMRES reserves 101 basis vectors, reference gmres.hh:31-32)
// only occupies device memory for the vectors a solve touches.
#ifndef FLECSOLVE_AMD_VECTORS_DEVICE_HH
#define FLECSOLVE_AMD_VECTORS_DEVICE_HH

#include <array>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

#include "fsx.h"
#include "flecsolve/vectors/core.hh"

namespace flecsolve {

// Throws std::runtime_error carrying fsx_last_error() on a failed call.
inline void fsx_check(int rc, const char * what) {
	if (rc != FSX_OK)
		throw std::runtime_error(std::string(what) + ": " + fsx_last_error());
}

}

namespace flecsolve::vec {

template<class T>
struct device_config {
	using scalar = T;
	using real = T;
	using len_t = std::size_t;
	using var_t = anon_var;
	static constexpr auto var = variable<anon_var::anonymous>;
	static constexpr std::size_t num_components = 1;
	static constexpr bool is_complex = false;
	static constexpr fsx_dtype dtype = std::is_same_v<T, double> ? FSX_F64 : FSX_F32;
};

template<class Config>
struct device_data {
	using config = Config;
	using scalar = typename Config::scalar;

	struct state {
		fsx_ctx * ctx;
		std::size_t n;
		fsx_vec * v = nullptr;
		std::uint64_t key;
		~state() {
			if (v) fsx_vec_release(v);
		}
	};

	device_data() = default;
	device_data(fsx_ctx * ctx, std::size_t n) : s(std::make_shared<state>()) {
		s->ctx = ctx;
		s->n = n;
		s->key = next_key();
	}
	// alias an existing library vector (takes a reference)
	static device_data borrow(fsx_ctx * ctx, fsx_vec * v) {
		device_data d;
		d.s = std::make_shared<state>();
		d.s->ctx = ctx;
		fsx_check(fsx_vec_size(v, &d.s->n), "fsx_vec_size");
		fsx_check(fsx_vec_retain(v), "fsx_vec_retain");
		d.s->v = v;
		d.s->key = next_key();
		return d;
	}

	fsx_ctx * ctx() const { return s->ctx; }
	std::size_t size() const { return s->n; }
	// a new (lazily allocated) buffer of the same length on the same context
	device_data make_like() const { return device_data(s->ctx, s->n); }
	fsx_vec * get() const {
		if (!s->v)
			fsx_check(fsx_vec_create(s->ctx, s->n, Config::dtype, &s->v),
			          "fsx_vec_create");
		return s->v;
	}
	friend bool operator==(const device_data & a, const device_data & b) {
		return a.s == b.s;
	}
	friend bool operator!=(const device_data & a, const device_data & b) {
		return !(a == b);
	}

private:
	static std::uint64_t next_key() {
		static std::uint64_t k = 0;
		return ++k;
	}
	std::shared_ptr<state> s;
};

// Value already resident on the host (the device reduction synchronised).
template<class T>
struct ready {
	T val;
	T get() const { return val; }
	void wait() const {}
};

template<class Data>
struct device_ops {
	using config = typename Data::config;
	using scalar = typename config::scalar;
	using real = typename config::real;
	using len_t = typename config::len_t;

	static void copy(const Data & s, Data & d) {
		fsx_check(fsx_copy(d.ctx(), s.get(), d.get()), "copy");
	}
	static void zero(Data & x) { fsx_check(fsx_zero(x.ctx(), x.get()), "zero"); }
	static void set_to_scalar(scalar a, Data & x) {
		fsx_check(fsx_set_scalar(x.ctx(), a, x.get()), "set_scalar");
	}
	static void scale(scalar a, Data & x) {
		fsx_check(fsx_scale(x.ctx(), a, x.get()), "scale");
	}
	static void scale(scalar a, const Data & x, Data & y) {
		fsx_check(fsx_scale_to(y.ctx(), a, x.get(), y.get()), "scale");
	}
	static void add(const Data & x, const Data & y, Data & z) {
		fsx_check(fsx_add(z.ctx(), x.get(), y.get(), z.get()), "add");
	}
	static void subtract(const Data & x, const Data & y, Data & z) {
		fsx_check(fsx_subtract(z.ctx(), x.get(), y.get(), z.get()), "subtract");
	}
	static void multiply(const Data & x, const Data & y, Data & z) {
		fsx_check(fsx_multiply(z.ctx(), x.get(), y.get(), z.get()), "multiply");
	}
	static void divide(const Data & x, const Data & y, Data & z) {
		fsx_check(fsx_divide(z.ctx(), x.get(), y.get(), z.get()), "divide");
	}
	static void reciprocal(const Data & x, Data & y) {
		fsx_check(fsx_reciprocal(y.ctx(), x.get(), y.get()), "reciprocal");
	}
	static void linear_sum(scalar a, const Data & x, scalar b, const Data & y,
	                       Data & z) {
		fsx_check(fsx_linear_sum(z.ctx(), a, x.get(), b, y.get(), z.get()),
		          "linear_sum");
	}
	static void axpy(scalar a, const Data & x, const Data & y, Data & z) {
		fsx_check(fsx_axpy(z.ctx(), a, x.get(), y.get(), z.get()), "axpy");
	}
	static void axpby(scalar a, scalar b, const Data & x, Data & z) {
		fsx_check(fsx_axpby(z.ctx(), a, b, x.get(), z.get()), "axpby");
	}
	static void abs(const Data & x, Data & y) {
		fsx_check(fsx_abs(y.ctx(), x.get(), y.get()), "abs");
	}
	static void add_scalar(const Data & x, scalar a, Data & y) {
		fsx_check(fsx_add_scalar(y.ctx(), x.get(), a, y.get()), "add_scalar");
	}
	static void set_random(Data & x, unsigned seed) {
		fsx_check(fsx_set_random(x.ctx(), x.get(), seed), "set_random");
	}
	// one value per line to "<pre>-<colour>" (the dump task of
	// vectors/operations/topo_tasks.hh:316-323; one colour here: "-0")
	static void dump(std::string_view pre, const Data & x) {
		std::vector<scalar> h(x.size());
		if (!h.empty())
			fsx_check(fsx_vec_download(x.ctx(), x.get(), h.data(), h.size()), "dump");
		std::string fname{pre};
		fname += "-0";
		std::ofstream ofile(fname);
		for (const scalar & v : h) ofile << v << '\n';
	}

	static ready<real> min(const Data & x) {
		double v;
		fsx_check(fsx_min(x.ctx(), x.get(), &v), "min");
		return {(real)v};
	}
	static ready<real> max(const Data & x) {
		double v;
		fsx_check(fsx_max(x.ctx(), x.get(), &v), "max");
		return {(real)v};
	}
	// p >= 3: (sum pow(x_i, p))^(1/p) as the FleCSI vector path computes it
	// (vectors/operations/topo_view.hh:243-257; no absolute value)
	template<unsigned short p>
	static ready<real> lp_norm(const Data & x) {
		static_assert(p >= 1, "lp_norm: p must be >= 1");
		double v;
		if constexpr (p == 1) fsx_check(fsx_nrm1(x.ctx(), x.get(), &v), "l1norm");
		else if constexpr (p == 2) fsx_check(fsx_nrm2(x.ctx(), x.get(), &v), "l2norm");
		else fsx_check(fsx_nrmp(x.ctx(), x.get(), p, &v), "lp_norm");
		return {(real)v};
	}
	static ready<real> inf_norm(const Data & x) {
		double v;
		fsx_check(fsx_nrminf(x.ctx(), x.get(), &v), "inf_norm");
		return {(real)v};
	}
	static ready<scalar> dot(const Data & x, const Data & y) {
		double v;
		fsx_check(fsx_dot(x.ctx(), x.get(), y.get(), &v), "dot");
		return {(scalar)v};
	}
	static ready<len_t> global_size(const Data & x) { return {x.size()}; }
	static len_t local_size(const Data & x) { return x.size(); }

	// One device element (vectors/seq.hh:271-280 returns a reference into
	// host storage): reading converts to scalar, assigning writes it; each
	// access is one small synchronous copy, ordered after queued work.
	struct element {
		const Data * d;
		len_t i;
		operator scalar() const {
			double v;
			fsx_check(fsx_vec_get(d->ctx(), d->get(), i, &v), "element read");
			return (scalar)v;
		}
		element & operator=(scalar a) {
			fsx_check(fsx_vec_set(d->ctx(), d->get(), i, (double)a), "element write");
			return *this;
		}
		element & operator=(const element & o) { return *this = (scalar)o; }
		element & operator+=(scalar a) { return *this = (scalar)*this + a; }
		element & operator-=(scalar a) { return *this = (scalar)*this - a; }
		element & operator*=(scalar a) { return *this = (scalar)*this * a; }
		element & operator/=(scalar a) { return *this = (scalar)*this / a; }
	};
	static element retreive(Data & x, len_t i) { return {&x, i}; }
	static scalar retreive(const Data & x, len_t i) { return (scalar)element{&x, i}; }
};

template<class T = double>
struct device : core<device_data, device_ops, device_config<T>> {
	using base = core<device_data, device_ops, device_config<T>>;
	device() : base{device_data<device_config<T>>{}} {}
	device(fsx_ctx * ctx, std::size_t n)
		: base{device_data<device_config<T>>(ctx, n)} {}
	explicit device(device_data<device_config<T>> d) : base{std::move(d)} {}
	// alias an fsx_vec owned elsewhere
	static device borrow(fsx_ctx * ctx, fsx_vec * v) {
		return device(device_data<device_config<T>>::borrow(ctx, v));
	}
	fsx_vec * handle() const { return this->data.get(); }
	fsx_ctx * context() const { return this->data.ctx(); }
};

// Work vectors shaped like `b` (the reference's work_factory,
// solvers/solver_settings.hh:319-330, creating FleCSI fields).
template<std::size_t nwork, class T>
std::array<device<T>, nwork> make_device_work(const device<T> & b) {
	std::array<device<T>, nwork> w;
	for (auto & v : w) v = device<T>(b.context(), b.data.size());
	return w;
}

}

#endif
