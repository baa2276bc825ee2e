// flecsolve/solvers/fused.hh -- CG, BiCGSTAB and GMRES on a device CSR as
// fused, graph-captured device solves.
//
// The generic solvers of cg.hh, bicgstab.hh and gmres.hh run the reference's
// iteration over any vec::core vector with one host round trip per
// reduction, as the reference does.  When the operator is a device CSR (and
// the preconditioner the identity or its Jacobi diagonal), the same
// iterations -- same operation order, tolerance tests and solve_info -- run
// as fused kernels with device-resident scalars captured in HIP graphs
// (fsx_cg_solve, fsx_bicgstab_solve, fsx_gmres_solve).  A bound solver
// (cg::solver(s, work)(op::ref(A), P)) hands its solves to them by itself
// (device_dispatch.hh); these policies are the same path held as an
// operator without a work array:
//
//   auto S = op::make_shared<cg::fused>(settings, A);        // or A, dinv
//   solve_info info = S.get().apply(b, x);
//
// ir::fused is the mixed-precision refinement (FP32 inner CG on the
// symmetrically Jacobi-scaled matrix) and ir::rr the FP32 CG with FP64
// residual replacement on that matrix, with the same interface.
//
// The matrix is held by handle: values changed in place (fsx_csr_set_values,
// fsx_csr_shift_scale) are seen by the next solve.
#ifndef FLECSOLVE_AMD_SOLVERS_FUSED_HH
#define FLECSOLVE_AMD_SOLVERS_FUSED_HH

#include <optional>
#include <type_traits>
#include <utility>

#include "fsx.h"
#include "flecsolve/matrices/device_csr.hh"
#include "flecsolve/solvers/bicgstab.hh"
#include "flecsolve/solvers/cg.hh"
#include "flecsolve/solvers/device_dispatch.hh"
#include "flecsolve/solvers/gmres.hh"

namespace flecsolve::detail {

using fused_fn = int (*)(fsx_ctx *, const fsx_csr *, const fsx_vec *, const fsx_vec *,
                         fsx_vec *, const fsx_settings *, fsx_info *);

template<class Settings, fused_fn Fn>
struct fused_solver : op::base<> {
	fused_solver(const Settings & s, mat::device_csr A) : settings(s), A(std::move(A)) {}
	fused_solver(const Settings & s, mat::device_csr A, const mat::device_dinv & D)
		: settings(s), A(std::move(A)), dinv(D.dinv) {}

	template<class B, class X>
	solve_info apply(const B & b, X & x) const {
		fsx_settings f = to_fsx(settings);
		if constexpr (std::is_same_v<Settings, gmres::settings>) {
			f.max_krylov_dim = settings.max_krylov_dim;
			f.pre_side = settings.pre_side == gmres::precond_side::left ? 0 : 1;
			f.restart = settings.restart;
		}
		fsx_info fi{};
		fsx_check(Fn(A.context(), A.handle(), dinv ? dinv->handle() : nullptr, b.data.get(),
		             x.data.get(), &f, &fi),
		          "fused solve");
		return from_fsx(fi);
	}

	Settings settings;
	mat::device_csr A;
	std::optional<vec::device<double>> dinv;
};

}

namespace flecsolve::cg {
using fused = detail::fused_solver<settings, fsx_cg_solve>;
}

namespace flecsolve::ir {

// FP32-inner / FP64-outer refinement (fsx_ir_cg_solve; no reference
// counterpart) as an operator: apply(b, x) refines x from its current value
// until |b - A x| / |b| <= settings.rtol (outer, FP64), each correction an
// FP32 CG on S A S (S = D^{-1/2}) to inner.rtol.  Holds an FP32 copy of A's
// values, written at construction; call refresh() after A's values change.
// solve_info: status, iters = inner iterations, restarts = outer
// iterations, res_norm_final = final relative residual.
struct fused : op::base<> {
	fused(const solver_settings & s, const solver_settings & inner, mat::device_csr A,
	      int max_outer = 20)
		: settings(s), inner(inner), A(std::move(A)), max_outer(max_outer) {
		fsx_csr * a32 = nullptr;
		fsx_check(fsx_csr_convert(this->A.context(), this->A.handle(), FSX_F32, &a32),
		          "ir: FP32 copy");
		A32 = mat::device_csr::adopt(this->A.context(), a32);
		refresh();
	}
	void refresh() {
		fsx_check(fsx_csr_set_values_f32_sym_scaled(A.context(), A32.handle(), A.handle()),
		          "ir: scaled FP32 values");
	}
	template<class B, class X>
	solve_info apply(const B & b, X & x) const {
		fsx_settings in = detail::to_fsx(inner);
		fsx_ir_info ii{};
		fsx_check(fsx_ir_cg_solve(A.context(), A.handle(), A32.handle(), 2, b.data.get(),
		                          x.data.get(), (double)settings.rtol, max_outer, &in, &ii),
		          "ir solve");
		solve_info info;
		info.status = static_cast<solve_info::stop_reason>(ii.status);
		info.iters = ii.inner_iters_total;
		info.restarts = ii.outer_iters;
		info.rhs_norm = (float)ii.rhs_norm;
		info.res_norm_final = (float)ii.rel_res_final;
		return info;
	}
	solver_settings settings, inner;
	mat::device_csr A, A32;
	int max_outer;
};

// FP32 CG with FP64 residual replacement (fsx_ir_cg_rr_solve; no reference
// counterpart) as an operator, with ir::fused's interface: apply(b, x)
// solves from x's current value until |b - A x| / |b| <= settings.rtol
// (FP64), one FP32 CG on S A S whose recursive residual is replaced by the
// true FP64 one after each segment of inner.rtol reduction (e.g. 3e-2; at
// most inner.maxiter iterations a segment), the search direction kept.
// solve_info: status, iters = FP32 iterations, restarts = segments,
// res_norm_final = final relative residual.
struct rr : op::base<> {
	rr(const solver_settings & s, const solver_settings & inner, mat::device_csr A,
	   int max_segments = 100)
		: settings(s), inner(inner), A(std::move(A)), max_segments(max_segments) {
		fsx_csr * a32 = nullptr;
		fsx_check(fsx_csr_convert(this->A.context(), this->A.handle(), FSX_F32, &a32),
		          "rr: FP32 copy");
		A32 = mat::device_csr::adopt(this->A.context(), a32);
		refresh();
	}
	void refresh() {
		fsx_check(fsx_csr_set_values_f32_sym_scaled(A.context(), A32.handle(), A.handle()),
		          "rr: scaled FP32 values");
	}
	template<class B, class X>
	solve_info apply(const B & b, X & x) const {
		fsx_settings in = detail::to_fsx(inner);
		fsx_ir_info ii{};
		fsx_check(fsx_ir_cg_rr_solve(A.context(), A.handle(), A32.handle(), b.data.get(),
		                             x.data.get(), (double)settings.rtol, max_segments, &in,
		                             &ii),
		          "rr solve");
		solve_info info;
		info.status = static_cast<solve_info::stop_reason>(ii.status);
		info.iters = ii.inner_iters_total;
		info.restarts = ii.outer_iters;
		info.rhs_norm = (float)ii.rhs_norm;
		info.res_norm_final = (float)ii.rel_res_final;
		return info;
	}
	solver_settings settings, inner;
	mat::device_csr A, A32;
	int max_segments;
};

}
namespace flecsolve::bicgstab {
using fused = detail::fused_solver<settings, fsx_bicgstab_solve>;
}
namespace flecsolve::gmres {
using fused = detail::fused_solver<settings, fsx_gmres_solve>;
}

#endif
