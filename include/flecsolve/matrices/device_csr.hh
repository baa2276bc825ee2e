// flecsolve/matrices/device_csr.hh -- the device CSR operator.
//
// Drop-in for the reference's sparse-matrix operators mat::csr /
// mat::parcsr (flecsolve/matrices/seq.hh:35-63, 257-261;
// flecsolve/matrices/parcsr.hh:100-177): an op::base policy whose apply(x, y)
// is y = A x.  The reference computes that as FleCSI tasks (spmv_remote on the
// ghosted off-diagonal block, spmv_local on the diagonal block, then y.add,
// parcsr.hh:60-68); on one GPU the off-diagonal block is empty and the
// product is a single libfsx SpMV kernel.  Use as op::core<mat::device_csr>.
#ifndef FLECSOLVE_AMD_MATRICES_DEVICE_CSR_HH
#define FLECSOLVE_AMD_MATRICES_DEVICE_CSR_HH

#include <memory>

#include "fsx.h"
#include "flecsolve/operators/core.hh"
#include "flecsolve/vectors/device.hh"

namespace flecsolve::mat {

struct device_csr : op::base<> {
	device_csr() = default;
	// borrow: the caller keeps ownership of A
	device_csr(fsx_ctx * c, fsx_csr * a) : ctx(c), A(a, [](fsx_csr *) {}) {}
	// adopt: destroyed with the last copy
	static device_csr adopt(fsx_ctx * c, fsx_csr * a) {
		device_csr m;
		m.ctx = c;
		m.A = std::shared_ptr<fsx_csr>(a, [](fsx_csr * p) { fsx_csr_destroy(p); });
		return m;
	}

	template<class D, class R>
	void apply(const D & x, R & y) const {
		fsx_check(fsx_csr_spmv(ctx, A.get(), x.data.get(), y.data.get()), "spmv");
	}
	template<class D, class R>
	void mult(const D & x, R & y) const { apply(x, y); }

	int64_t rows() const { int64_t r; fsx_csr_info(A.get(), &r, nullptr, nullptr, nullptr, nullptr); return r; }
	int64_t cols() const { int64_t c; fsx_csr_info(A.get(), nullptr, &c, nullptr, nullptr, nullptr); return c; }
	int64_t nnz() const { int64_t z; fsx_csr_info(A.get(), nullptr, nullptr, &z, nullptr, nullptr); return z; }
	fsx_csr * handle() const { return A.get(); }
	fsx_ctx * context() const { return ctx; }

	fsx_ctx * ctx = nullptr;
	std::shared_ptr<fsx_csr> A;
};

// Diagonal-inverse preconditioner z = D^{-1} r: the Dinv CSR operator of the
// reference's tests (flecsolve/util/test/mesh.hh:123-140) held as a vector.
struct device_dinv : op::base<> {
	device_dinv() = default;
	explicit device_dinv(const device_csr & A)
		: dinv(A.context(), (std::size_t)A.rows()) {
		fsx_check(fsx_csr_diag_inv(A.context(), A.handle(), dinv.handle()),
		          "diag_inv");
	}
	template<class D, class R>
	void apply(const D & r, R & z) const {
		fsx_check(fsx_jacobi_apply(z.data.ctx(), dinv.handle(), r.data.get(),
		                           z.data.get()),
		          "jacobi_apply");
	}
	vec::device<double> dinv;
};

}

namespace flecsolve::mg {

// Weighted Jacobi relaxation as an operator (mg::bound_jacobi, reference
// flecsolve/solvers/mg/jacobi.hh:25-53, 98-113): apply(b, x) runs `nrelax`
// sweeps of x = omega D^{-1}(b - (L+U) x) + (1 - omega) x in place.
struct jacobi_settings {
	float omega = 2 / 3.;
	std::size_t nrelax = 1;
};

struct bound_jacobi : op::base<> {
	bound_jacobi(mat::device_csr a, jacobi_settings s) : A(std::move(a)), settings(s) {}
	template<class D, class R>
	void apply(const D & b, R & x) const {
		fsx_check(fsx_jacobi_relax(A.context(), A.handle(), settings.omega,
		                           (int)settings.nrelax, b.data.get(), x.data.get()),
		          "jacobi_relax");
	}
	mat::device_csr A;
	jacobi_settings settings;
};

}

#endif

// the bound Krylov solvers' hand-off to the fused drivers for this operator
#include "flecsolve/solvers/device_dispatch.hh"
