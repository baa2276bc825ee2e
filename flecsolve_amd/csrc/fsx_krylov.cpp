// fsx_krylov.cpp -- C ABI entry points that run the generic C++ solvers
// (include/flecsolve/solvers/{cg,gmres,bicgstab}.hh) over the device vector
// type vec::device<double> and the device CSR operator -- the drop-in path in
// which the solver code is the reference's algorithm unchanged and every
// vector operation is a libfsx kernel.
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>

#include "fsx.h"
#include "flecsolve/matrices/device_csr.hh"
#include "flecsolve/solvers/bicgstab.hh"
#include "flecsolve/solvers/cg.hh"
#include "flecsolve/solvers/gmres.hh"
#include "flecsolve/solvers/mg/ua.hh"
#include "flecsolve/vectors/device.hh"

namespace fsx {
void set_error(const std::string & msg);
}

namespace {

using namespace flecsolve;
using dvec = vec::device<double>;

struct history {
	double * out;
	int cap;
	int n = 0;
	template<class V>
	bool operator()(const V &, double res) {
		if (out && n < cap) out[n] = res;
		++n;
		return false;
	}
};

void to_info(const solve_info & s, fsx_info * o) {
	o->status = (int)s.status;
	o->iters = s.iters;
	o->restarts = s.restarts;
	o->res_norm_initial = s.res_norm_initial;
	o->res_norm_final = s.res_norm_final;
	o->sol_norm_initial = s.sol_norm_initial;
	o->sol_norm_final = s.sol_norm_final;
	o->rhs_norm = s.rhs_norm;
}

template<class S>
S to_settings(const fsx_settings * s) {
	S o;
	o.maxiter = s->maxiter;
	o.rtol = s->rtol;
	o.atol = s->atol;
	o.use_zero_guess = s->use_zero_guess != 0;
	return o;
}

template<class F>
int guarded(F && f) {
	try {
		f();
		return FSX_OK;
	}
	catch (const std::invalid_argument & e) {
		fsx::set_error(e.what());
		return FSX_ERR_ARG;
	}
	catch (const std::exception & e) {
		fsx::set_error(e.what());
		return FSX_ERR_STATE;
	}
}

// Run `solve(A, P)` with the requested preconditioner handle.
template<class Solve>
void with_prec(fsx_ctx * ctx, const fsx_csr * A, int prec, int nrelax,
               Solve && solve) {
	op::core<mat::device_csr> Aop(ctx, const_cast<fsx_csr *>(A));
	if (prec == 0) solve(op::ref(Aop), op::I);
	else if (prec == 1) {
		op::core<mat::device_dinv> D(Aop);
		solve(op::ref(Aop), op::ref(D));
	}
	else if (prec == 2) {
		mg::jacobi_settings js;
		js.nrelax = (std::size_t)nrelax;
		op::core<mg::bound_jacobi> J(static_cast<const mat::device_csr &>(Aop), js);
		solve(op::ref(Aop), op::ref(J));
	}
	else
		throw std::invalid_argument("prec must be 0 (identity), 1 (Jacobi) or 2 "
		                            "(weighted Jacobi relaxation)");
}

void check_args(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * b,
                const fsx_vec * x, const fsx_settings * s, fsx_info * info) {
	if (!ctx || !A || !b || !x || !s || !info)
		throw std::invalid_argument("null argument");
	int64_t nr, nc;
	int vt;
	fsx_csr_info(A, &nr, &nc, nullptr, nullptr, &vt);
	size_t nb, nx;
	fsx_vec_size(b, &nb);
	fsx_vec_size(x, &nx);
	int db, dx;
	fsx_vec_dtype(b, &db);
	fsx_vec_dtype(x, &dx);
	if (nr != nc || (int64_t)nb != nr || (int64_t)nx != nr)
		throw std::invalid_argument("shape mismatch");
	if (db != FSX_F64 || dx != FSX_F64 || vt != FSX_F64)
		throw std::invalid_argument("generic solvers run in FP64");
}

} // namespace

extern "C" {

int fsx_krylov_cg(fsx_ctx * ctx, const fsx_csr * A, int prec, int nrelax,
                  const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                  fsx_info * info, double * res_hist, int hist_cap) {
	return guarded([&] {
		check_args(ctx, A, b, x, s, info);
		dvec bv = dvec::borrow(ctx, const_cast<fsx_vec *>(b));
		dvec xv = dvec::borrow(ctx, x);
		history h{res_hist, hist_cap};
		with_prec(ctx, A, prec, nrelax, [&](auto Ah, auto Ph) {
			cg::solver slv(to_settings<cg::settings>(s),
			               vec::make_device_work<cg::nwork>(bv));
			// the history is a diagnostic only when one is asked for (a
			// bound diagnostic is called after every iteration)
			if (res_hist && hist_cap > 0) to_info(slv(Ah, Ph, std::ref(h))(bv, xv), info);
			else to_info(slv(Ah, Ph)(bv, xv), info);
		});
	});
}

int fsx_krylov_cg_generic(fsx_ctx * ctx, const fsx_csr * A, int prec, int nrelax,
                          const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                          fsx_info * info, double * res_hist, int hist_cap) {
	return guarded([&] {
		check_args(ctx, A, b, x, s, info);
		dvec bv = dvec::borrow(ctx, const_cast<fsx_vec *>(b));
		dvec xv = dvec::borrow(ctx, x);
		history h{res_hist, hist_cap};
		with_prec(ctx, A, prec, nrelax, [&](auto Ah, auto Ph) {
			cg::solver slv(to_settings<cg::settings>(s),
			               vec::make_device_work<cg::nwork>(bv));
			to_info(slv(Ah, Ph, std::ref(h)).generic(bv, xv), info);
		});
	});
}

int fsx_krylov_fcg(fsx_ctx * ctx, const fsx_csr * A, int prec, int nrelax,
                   const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                   fsx_info * info) {
	return guarded([&] {
		check_args(ctx, A, b, x, s, info);
		dvec bv = dvec::borrow(ctx, const_cast<fsx_vec *>(b));
		dvec xv = dvec::borrow(ctx, x);
		with_prec(ctx, A, prec, nrelax, [&](auto Ah, auto Ph) {
			fcg::solver slv(to_settings<fcg::settings>(s),
			                vec::make_device_work<fcg::nwork>(bv));
			auto op = slv(Ah, Ph);
			to_info(op(bv, xv), info);
		});
	});
}

int fsx_krylov_gmres(fsx_ctx * ctx, const fsx_csr * A, int prec, int nrelax,
                     const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                     fsx_info * info, double * res_hist, int hist_cap) {
	return guarded([&] {
		check_args(ctx, A, b, x, s, info);
		dvec bv = dvec::borrow(ctx, const_cast<fsx_vec *>(b));
		dvec xv = dvec::borrow(ctx, x);
		history h{res_hist, hist_cap};
		gmres::settings gs = to_settings<gmres::settings>(s);
		gs.max_krylov_dim = s->max_krylov_dim;
		gs.pre_side = s->pre_side == 0 ? gmres::precond_side::left
		                               : gmres::precond_side::right;
		gs.restart = s->restart != 0;
		with_prec(ctx, A, prec, nrelax, [&](auto Ah, auto Ph) {
			gmres::solver slv(gs, vec::make_device_work<gmres::nwork>(bv));
			auto op = slv(Ah, Ph, std::ref(h));
			to_info(op(bv, xv), info);
		});
	});
}

int fsx_krylov_bicgstab(fsx_ctx * ctx, const fsx_csr * A, int prec,
                        int nrelax, const fsx_vec * b, fsx_vec * x,
                        const fsx_settings * s, fsx_info * info,
                        double * res_hist, int hist_cap) {
	return guarded([&] {
		check_args(ctx, A, b, x, s, info);
		dvec bv = dvec::borrow(ctx, const_cast<fsx_vec *>(b));
		dvec xv = dvec::borrow(ctx, x);
		history h{res_hist, hist_cap};
		with_prec(ctx, A, prec, nrelax, [&](auto Ah, auto Ph) {
			bicgstab::solver slv(to_settings<bicgstab::settings>(s),
			                     vec::make_device_work<bicgstab::nwork>(bv));
			auto op = slv(Ah, Ph, std::ref(h));
			to_info(op(bv, xv), info);
		});
	});
}

// CG / FCG preconditioned by an AMG solver (mg::ua::bound_solver as P).
int fsx_krylov_cg_amg(fsx_ctx * ctx, const fsx_csr * A, fsx_amg * M, int flexible,
                      const fsx_vec * b, fsx_vec * x, const fsx_settings * s, fsx_info * info,
                      double * res_hist, int hist_cap) {
	return guarded([&] {
		check_args(ctx, A, b, x, s, info);
		if (!M) throw std::invalid_argument("null AMG solver");
		dvec bv = dvec::borrow(ctx, const_cast<fsx_vec *>(b));
		dvec xv = dvec::borrow(ctx, x);
		history h{res_hist, hist_cap};
		op::core<mat::device_csr> Aop(ctx, const_cast<fsx_csr *>(A));
		op::core<mg::ua::bound_solver> Pm(static_cast<const mat::device_csr &>(Aop), M);
		if (flexible) {
			fcg::solver slv(to_settings<cg::settings>(s), vec::make_device_work<fcg::nwork>(bv));
			auto op = slv(op::ref(Aop), op::ref(Pm), std::ref(h));
			to_info(op(bv, xv), info);
		}
		else {
			cg::solver slv(to_settings<cg::settings>(s), vec::make_device_work<cg::nwork>(bv));
			auto op = slv(op::ref(Aop), op::ref(Pm), std::ref(h));
			to_info(op(bv, xv), info);
		}
	});
}

} // extern "C"
