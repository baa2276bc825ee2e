// flecsolve/time-integrators/operator_adapter.hh -- the implicit stage
// operator I - gamma f built from a right-hand-side operator f.
//
// Interface of the reference's
// flecsolve/time-integrators/operator_adapter.hh:24-50: operator_adapter<Op>
// derives from Op; apply_rhs(x, y) is Op's apply (y = f(x)); apply(x, y) is
// y = x - gamma f(x), computed as f(x) then y.axpy(-gamma, y, x); the
// integrator sets gamma with set_scaling before each solve.
//
// csr_adapter is the device-native form for a right-hand side that is a
// sparse matrix F: it keeps a second CSR holding I - gamma F (same pattern,
// values rewritten on the device by one kernel when gamma changes) so that a
// solve of the stage system runs as one matrix -- for instance through the
// fused, graph-captured cg::fused solver -- instead of an SpMV and an axpy
// per product.
#ifndef FLECSOLVE_AMD_TIME_INTEGRATORS_OPERATOR_ADAPTER_HH
#define FLECSOLVE_AMD_TIME_INTEGRATORS_OPERATOR_ADAPTER_HH

#include <memory>
#include <utility>

#include "flecsolve/matrices/device_csr.hh"

namespace flecsolve::time_integrator {

template<class Op>
struct operator_adapter : Op {
	template<class... Args>
	operator_adapter(Args &&... args) : Op(std::forward<Args>(args)...), gamma{1.} {}

	template<class D, class R>
	void apply(const D & x, R & y) const {
		apply_rhs(x, y);
		y.axpy(-gamma, y, x);
	}

	template<class D, class R>
	decltype(auto) apply_rhs(const D & x, R & y) const {
		return static_cast<const Op &>(*this).apply(x, y);
	}

	template<class V>
	bool is_valid(const V &) { return true; }

	double get_scaling() const { return gamma; }
	void set_scaling(double scaling) { gamma = scaling; }

protected:
	double gamma;
};

// f(x) = F x for a device CSR F; the stage operator is the CSR I - gamma F.
// F must be square, FP64, and store every row's diagonal entry (the identity
// is added there; construction throws otherwise).
struct csr_adapter : op::base<> {
	explicit csr_adapter(mat::device_csr F) : F(std::move(F)) {
		fsx_csr * s = nullptr;
		fsx_check(fsx_csr_convert(this->F.context(), this->F.handle(), FSX_F64, &s),
		          "csr_adapter: stage matrix");
		stage = mat::device_csr::adopt(this->F.context(), s);
		set_scaling(1.0);
	}

	// y = F x
	template<class D, class R>
	void apply_rhs(const D & x, R & y) const { F.apply(x, y); }
	// y = (I - gamma F) x
	template<class D, class R>
	void apply(const D & x, R & y) const { stage.apply(x, y); }

	template<class V>
	bool is_valid(const V &) { return true; }

	double get_scaling() const { return gamma; }
	void set_scaling(double g) {
		gamma = g;
		fsx_check(fsx_csr_shift_scale(F.context(), F.handle(), 1.0, -g, stage.handle()),
		          "csr_adapter: set_scaling");
	}

	// the CSR holding I - gamma F (for solvers that take a matrix)
	const mat::device_csr & matrix() const { return stage; }
	const mat::device_csr & rhs_matrix() const { return F; }

protected:
	mat::device_csr F, stage;
	double gamma = 1.0;
};

}

#endif
