// C++ drop-in check: the reference's CG test pattern
// (flecsolve/solvers/test/cg.cc:16-91) written against the device vector
// type and device CSR operator.  b = 0, x0 = set_random(7), and a user
// diagnostic that tracks the A-norm of the error (the solution is 0):
// monotone decrease and the Trefethen-Bau bound 2((sqrt k - 1)/(sqrt k + 1))^i.
// Also runs GMRES and BiCGSTAB through the same API, and CG on a Matrix
// Market file read with matrix_market<>::read(...).tocsr() (argv[1]), and
// solvers chosen at run time from config files (argv[2]: tests/golden), and
// the fused CG and mixed-precision refinement policies (solvers/fused.hh),
// and CG over the parallel CSR operator (matrices/device_parcsr.hh).
// Exit status 0 = pass.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <string>
#include <vector>

#include "flecsolve/matrices/device_csr.hh"
#include "flecsolve/matrices/device_parcsr.hh"
#include "flecsolve/matrices/io/matrix_market.hh"
#include "flecsolve/solvers/bicgstab.hh"
#include "flecsolve/solvers/cg.hh"
#include "flecsolve/solvers/factory.hh"
#include "flecsolve/solvers/fused.hh"
#include "flecsolve/solvers/gmres.hh"
#include "flecsolve/solvers/mg/gs.hh"
#include "fsx_problems.h"

using namespace flecsolve;

// z = one hybrid Gauss-Seidel application to r from z = 0
struct sgs_precond : op::base<> {
	explicit sgs_precond(op::core<op::hybrid_gs> g) : G(std::move(g)) {}
	template<class D, class R>
	void apply(const D & r, R & z) const {
		z.set_scalar(0.);
		G.apply(r, z);
	}
	op::core<op::hybrid_gs> G;
};

template<class Op, class Vec>
struct diagnostic {
	diagnostic(const Op & A, const Vec & x0, double cond)
		: cond(cond), A(A), Ax(x0.context(), x0.data.size()) {
		A.apply(x0, Ax);
		e0 = std::sqrt(x0.dot(Ax).get());
		eprev = e0;
	}
	template<class T>
	bool operator()(const T & x, double) {
		A.apply(x, Ax);
		const double ea = std::sqrt(x.dot(Ax).get());
		const double frac = (std::sqrt(cond) - 1) / (std::sqrt(cond) + 1);
		const double bnd = 2 * std::pow(frac, ++iter) * e0;
		monotonic_fail = monotonic_fail || ea > eprev;
		convergence_fail = convergence_fail || ea >= bnd;
		eprev = ea;
		return false;
	}
	int iter = 0;
	double cond, e0, eprev;
	const Op & A;
	Vec Ax;
	bool monotonic_fail = false, convergence_fail = false;
};

int main(int argc, char ** argv) {
	const int n = 48;
	std::vector<int32_t> rp(n * n + 1), ci(fsx_gen_poisson2d_nnz(n));
	std::vector<double> v(ci.size());
	fsx_gen_poisson2d(n, rp.data(), ci.data(), v.data());
	double lmin = 1e300, lmax = 0;
	for (int i = 1; i <= n; ++i)
		for (int j = 1; j <= n; ++j) {
			double l = 4 - 2 * std::cos(M_PI * i / (n + 1)) - 2 * std::cos(M_PI * j / (n + 1));
			lmin = std::min(lmin, l);
			lmax = std::max(lmax, l);
		}
	fsx_ctx * ctx;
	fsx_check(fsx_ctx_create(0, &ctx), "ctx");
	fsx_csr * Ah;
	fsx_check(fsx_csr_create(ctx, n * n, n * n, (int64_t)ci.size(), FSX_I32, rp.data(),
	                         ci.data(), FSX_F64, v.data(), &Ah),
	          "csr");
	int fails = 0;
	{
		// the rest of the vector surface: lp_norm<3> of the FleCSI path
		// ((sum x^3)^(1/3), no abs: 7^(1/3) for [-1, 2]) and dump to
		// "<pre>-0", one value per line (topo_tasks.hh:216-226, 316-323)
		vec::device<double> w(ctx, 2);
		const double hw[2] = {-1.0, 2.0};
		fsx_check(fsx_vec_upload(ctx, w.data.get(), hw, 2), "upload");
		const double l3 = w.lp_norm<3>().get();
		const std::string pre = std::string(std::getenv("TMPDIR") ? std::getenv("TMPDIR") : "/tmp") +
		                        "/fsx_dump_check";
		w.dump(pre);
		std::ifstream in(pre + "-0");
		double a = 0, b = 0;
		in >> a >> b;
		std::printf("vector surface: lp_norm<3> %.17g, dump read back %g %g\n", l3, a, b);
		fails += std::abs(l3 - std::cbrt(7.0)) > 1e-15 || a != -1.0 || b != 2.0;
		std::remove((pre + "-0").c_str());
		// element access, operator[] / Ops::retreive (core.hh:338-340)
		w[0] = 5.0;
		w[1] += 0.5;
		const auto & cw = w;
		const double e0 = w[0], e1 = cw[1];
		std::printf("vector surface: w[0] = %g, w[1] = %g after w[0] = 5, w[1] += 0.5\n", e0, e1);
		fails += e0 != 5.0 || e1 != 2.5;
	}
	{
		op::core<mat::device_csr> A(ctx, Ah);
		vec::device<double> b(ctx, n * n), x(ctx, n * n);
		b.set_scalar(0.0);
		x.set_random(7);
		diagnostic diag{A, x, lmax / lmin};
		cg::settings s;
		s.maxiter = 2000;
		s.rtol = 1e-9;
		s.use_zero_guess = false;
		auto slv = cg::solver(s, vec::make_device_work<cg::nwork>(b))(
			op::ref(A), op::I, std::ref(diag));
		solve_info info = slv(b, x);
		std::printf("cg: %d iterations, %s, monotone %s, bound %s\n", info.iters,
		            info.success() ? "converged" : "not converged",
		            diag.monotonic_fail ? "FAIL" : "ok",
		            diag.convergence_fail ? "FAIL" : "ok");
		fails += !info.success() || diag.monotonic_fail || diag.convergence_fail ||
		         info.iters != diag.iter;

		// GMRES(30) with the Dinv preconditioner, BiCGSTAB, from x0 = 0
		vec::device<double> rhs(ctx, n * n), y(ctx, n * n);
		rhs.set_random(0);
		gmres::settings gs;
		gs.maxiter = 3000; // restarted GMRES(30) on the 48^2 Laplacian needs ~450
		gs.rtol = 1e-10;
		gs.use_zero_guess = true;
		gs.max_krylov_dim = 30;
		gs.restart = true;
		op::core<mat::device_dinv> D(A);
		auto gm = gmres::solver(gs, vec::make_device_work<gmres::nwork>(rhs))(
			op::ref(A), op::ref(D));
		solve_info gi = gm(rhs, y);
		vec::device<double> r(ctx, n * n);
		A.residual(rhs, y, r);
		const double grel = r.l2norm().get() / rhs.l2norm().get();
		std::printf("gmres(30)+dinv: %d iterations, %d restarts, true rel res %.3e\n",
		            gi.iters, gi.restarts, grel);
		fails += !gi.success() || grel > 1e-9;
		bicgstab::settings bs;
		bs.maxiter = 400;
		bs.rtol = 1e-10;
		bs.use_zero_guess = true;
		auto bi = bicgstab::solver(bs, vec::make_device_work<bicgstab::nwork>(rhs))(
			op::ref(A));
		solve_info bsi = bi(rhs, y);
		A.residual(rhs, y, r);
		const double brel = r.l2norm().get() / rhs.l2norm().get();
		std::printf("bicgstab: %d iterations, true rel res %.3e\n", bsi.iters, brel);
		fails += !bsi.success() || brel > 1e-9;

		// CG preconditioned by one symmetric hybrid Gauss-Seidel sweep from a
		// zero guess (a fixed SPD operator; the reference's multigrid solver
		// zeroes x the same way before cycling, solvers/mg/ua.hh:140-141),
		// built with mg::hybrid_gs::solver as in its levels (ua.hh:191-196)
		mg::hybrid_gs::solver gsf{{1, mg::relax_sweep::symmetric}};
		op::core<sgs_precond> SG(gsf(A));
		cg::settings fs;
		fs.maxiter = 500;
		fs.rtol = 1e-10;
		fs.use_zero_guess = true;
		auto fc = cg::solver(fs, vec::make_device_work<cg::nwork>(rhs))(op::ref(A),
		                                                               op::ref(SG));
		solve_info fi = fc(rhs, y);
		A.residual(rhs, y, r);
		const double frel = r.l2norm().get() / rhs.l2norm().get();
		auto pc = cg::solver(fs, vec::make_device_work<cg::nwork>(rhs))(op::ref(A));
		vec::device<double> y2(ctx, n * n);
		solve_info ci2 = pc(rhs, y2);
		std::printf("cg+symmetric hybrid_gs: %d iterations (no preconditioner: %d), "
		            "true rel res %.3e\n", fi.iters, ci2.iters, frel);
		fails += !fi.success() || frel > 1e-9 || fi.iters >= ci2.iters;
	}
	{
		// the reference-shaped call on the device operator hands the solve to
		// the fused driver: the same solve_info and x bits as calling
		// fsx_cg_solve / fsx_bicgstab_solve / fsx_gmres_solve directly, with a
		// host synchronisation per chunk of iterations instead of three per
		// iteration; generic() still runs the reference's sequence
		op::core<mat::device_csr> A(ctx, Ah);
		op::core<mat::device_dinv> D(A);
		const std::size_t m = (std::size_t)n * n;
		vec::device<double> rhs(ctx, m), y1(ctx, m), y2(ctx, m), y3(ctx, m);
		rhs.set_random(11);
		cg::settings cs;
		cs.maxiter = 3000;
		cs.rtol = 1e-10;
		cs.use_zero_guess = true;
		auto same = [&](const solve_info & a, const fsx_info & f, vec::device<double> & u,
		                vec::device<double> & w) {
			std::vector<double> hu(m), hw(m);
			fsx_check(fsx_vec_download(ctx, u.handle(), hu.data(), m), "download");
			fsx_check(fsx_vec_download(ctx, w.handle(), hw.data(), m), "download");
			return (int)a.status == f.status && a.iters == f.iters &&
			       a.restarts == f.restarts && a.res_norm_final == f.res_norm_final &&
			       a.sol_norm_final == f.sol_norm_final && a.rhs_norm == f.rhs_norm &&
			       hu == hw;
		};
		uint64_t s0, s1, s2, l;
		auto slv = cg::solver(cs, vec::make_device_work<cg::nwork>(rhs))(op::ref(A), op::ref(D));
		fsx_ctx_counters(ctx, &s0, &l);
		solve_info a = slv(rhs, y1);
		fsx_ctx_counters(ctx, &s1, &l);
		fsx_settings f = detail::to_fsx(cs);
		fsx_info fi{};
		fsx_check(fsx_cg_solve(ctx, Ah, D.dinv.handle(), rhs.handle(), y2.handle(), &f, &fi),
		          "cg");
		const bool cg_same = same(a, fi, y1, y2);
		s1 = s1 - s0;
		fsx_ctx_counters(ctx, &s0, &l);
		solve_info g = slv.generic(rhs, y3);
		fsx_ctx_counters(ctx, &s2, &l);
		s2 = s2 - s0;
		std::printf("dispatch cg+dinv: %d iterations = fsx_cg_solve: %s; host syncs: "
		            "reference-shaped call %llu, generic() %llu (%d iterations)\n",
		            a.iters, cg_same ? "yes" : "NO", (unsigned long long)s1,
		            (unsigned long long)s2, g.iters);
		fails += !a.success() || !cg_same || s1 * 4 > (uint64_t)a.iters ||
		         s2 < 3 * (uint64_t)g.iters || std::abs(g.iters - a.iters) > 2;
		bicgstab::settings bs;
		bs.maxiter = 400;
		bs.rtol = 1e-10;
		bs.use_zero_guess = true;
		auto bi = bicgstab::solver(bs, vec::make_device_work<bicgstab::nwork>(rhs))(op::ref(A));
		solve_info ba = bi(rhs, y1);
		f = detail::to_fsx(bs);
		fsx_check(fsx_bicgstab_solve(ctx, Ah, nullptr, rhs.handle(), y2.handle(), &f, &fi),
		          "bicgstab");
		const bool bi_same = same(ba, fi, y1, y2);
		gmres::settings gs;
		gs.maxiter = 3000;
		gs.rtol = 1e-10;
		gs.use_zero_guess = true;
		gs.max_krylov_dim = 30;
		gs.restart = true;
		auto gm = gmres::solver(gs, vec::make_device_work<gmres::nwork>(rhs))(op::ref(A),
		                                                                      op::ref(D));
		solve_info ga = gm(rhs, y1);
		f = detail::to_fsx(gs);
		f.max_krylov_dim = 30;
		f.pre_side = 1;
		f.restart = 1;
		fsx_check(fsx_gmres_solve(ctx, Ah, D.dinv.handle(), rhs.handle(), y2.handle(), &f, &fi),
		          "gmres");
		const bool gm_same = same(ga, fi, y1, y2);
		std::printf("dispatch bicgstab: %d iterations = fsx_bicgstab_solve: %s; gmres(30)+dinv: "
		            "%d iterations = fsx_gmres_solve: %s\n",
		            ba.iters, bi_same ? "yes" : "NO", ga.iters, gm_same ? "yes" : "NO");
		fails += !ba.success() || !bi_same || !ga.success() || !gm_same;
		// a user diagnostic: the fused per-iteration path, stopping the solve
		// at the iteration it asks for (converged_user, iters = it + 1)
		int calls = 0;
		auto stop7 = [&](const auto &, double) { return ++calls == 7; };
		auto ds = cg::solver(cs, vec::make_device_work<cg::nwork>(rhs))(op::ref(A), op::ref(D),
		                                                                 stop7);
		solve_info di = ds(rhs, y1);
		calls = 0;
		solve_info dg = ds.generic(rhs, y2);
		y1.axpy(-1.0, y2, y1);
		const double ddx = y1.l2norm().get() / y2.l2norm().get();
		std::printf("dispatch cg with a stopping diagnostic: %d iterations, %s; generic %d, "
		            "|dx| %.1e\n",
		            di.iters, di.status == solve_info::stop_reason::converged_user ? "user" : "?",
		            dg.iters, ddx);
		fails += di.iters != 7 || dg.iters != 7 ||
		         di.status != solve_info::stop_reason::converged_user ||
		         dg.status != solve_info::stop_reason::converged_user || ddx > 1e-12;
	}
	{
		// the fused solves behind the operator interface: CG + Dinv and the
		// mixed-precision refinement, against each other
		op::core<mat::device_csr> A(ctx, Ah);
		mat::device_csr Am(ctx, Ah);
		vec::device<double> rhs(ctx, n * n), y1(ctx, n * n), y2(ctx, n * n), r(ctx, n * n);
		rhs.set_random(5);
		cg::settings cs;
		cs.maxiter = 3000;
		cs.rtol = 1e-10;
		cs.use_zero_guess = true;
		mat::device_dinv D(Am);
		auto C = op::make_shared<cg::fused>(cs, Am, D);
		solve_info c1 = C.get().apply(rhs, y1);
		solver_settings inner;
		inner.maxiter = 3000;
		inner.rtol = 1e-4f;
		auto R = op::make_shared<ir::fused>(cs, inner, Am);
		y2.set_scalar(0.0);
		solve_info c2 = R.get().apply(rhs, y2);
		A.residual(rhs, y2, r);
		const double rel2 = r.l2norm().get() / rhs.l2norm().get();
		y2.axpy(-1.0, y1, y2);
		const double dx = y2.l2norm().get() / y1.l2norm().get();
		std::printf("fused cg+dinv: %d iterations; refinement: %d outer, %d inner, "
		            "true rel res %.3e, |x_ir - x_cg| / |x_cg| %.2e\n",
		            c1.iters, c2.restarts, c2.iters, rel2, dx);
		fails += !c1.success() || !c2.success() || rel2 > 1.01e-10 || dx > 1e-6;
		// FP32 CG with FP64 residual replacement behind the same interface
		vec::device<double> y3(ctx, n * n);
		solver_settings seg;
		seg.maxiter = 3000;
		seg.rtol = 3e-2f;
		auto RR = op::make_shared<ir::rr>(cs, seg, Am);
		y3.set_scalar(0.0);
		solve_info c3 = RR.get().apply(rhs, y3);
		A.residual(rhs, y3, r);
		const double rel3 = r.l2norm().get() / rhs.l2norm().get();
		y3.axpy(-1.0, y1, y3);
		const double dx3 = y3.l2norm().get() / y1.l2norm().get();
		std::printf("residual replacement: %d segments, %d iterations, true rel res %.3e, "
		            "|x_rr - x_cg| / |x_cg| %.2e\n",
		            c3.restarts, c3.iters, rel3, dx3);
		fails += !c3.success() || rel3 > 1.01e-10 || dx3 > 1e-6 || c3.iters > c2.iters;
	}
	{
		// the reference's parcsr test pattern (matrices/test/par_csr.cc:21-35):
		// b = 0, x = 2, CG through op::core of the parallel CSR; here over 1
		// and 4 colours against the plain CSR operator.  One colour is the
		// plain product, so its solve is the same bit for bit; four colours
		// round some rows differently and must converge alike.
		op::core<mat::device_csr> A(ctx, Ah);
		cg::settings ps;
		ps.maxiter = 3000;
		ps.rtol = 1e-8;
		ps.use_zero_guess = false;
		vec::device<double> pb(ctx, n * n), px(ctx, n * n), r(ctx, n * n);
		pb.set_scalar(0.0);
		px.set_scalar(2.0);
		auto plain = cg::solver(ps, vec::make_device_work<cg::nwork>(pb))(op::ref(A));
		// the parallel CSR runs the generic sequence: compare with the plain
		// operator's generic() (the reference-shaped call on the plain operator
		// takes the fused path and sums its dot products in another order)
		solve_info p0 = plain.generic(pb, px);
		int iters[2];
		double dx[2];
		for (int k = 0; k < 2; ++k) {
			op::core<mat::device_parcsr> M(ctx, Ah, k == 0 ? 1 : 4);
			vec::device<double> y(ctx, n * n);
			y.set_scalar(2.0);
			auto slv = cg::solver(ps, vec::make_device_work<cg::nwork>(pb))(op::ref(M));
			solve_info pi = slv(pb, y);
			iters[k] = pi.iters;
			y.axpy(-1.0, px, y);
			// the 2-norm: inf_norm starts from the smallest normal double
			// (vectors/seq.hh:251), so it never reads exactly 0
			dx[k] = y.l2norm().get();
			fails += !pi.success();
		}
		std::printf("parcsr cg (b = 0, x0 = 2): plain %d iterations; 1 colour %d (|dx| "
		            "%.1e); 4 colours %d (|dx| %.1e)\n",
		            p0.iters, iters[0], dx[0], iters[1], dx[1]);
		fails += !p0.success() || iters[0] != p0.iters || dx[0] != 0.0 ||
		         std::abs(iters[1] - p0.iters) > 2 || dx[1] > 1e-6;
	}
	if (argc > 2) {
		// run-time solver choice from config files (krylov_factory), as the
		// reference's applications build their linear solver
		op::core<mat::device_csr> A(ctx, Ah);
		vec::device<double> rhs(ctx, n * n), y(ctx, n * n), r(ctx, n * n);
		rhs.set_random(3);
		for (const char * name : {"krylov_cg.cfg", "krylov_gmres.cfg", "krylov_bicgstab.cfg"}) {
			const std::string path = std::string(argv[2]) + "/" + name;
			auto fs = read_config(path.c_str(), krylov_factory::options("linear-solver"));
			auto slv = krylov_factory::make(fs, rhs, op::ref(A));
			y.set_scalar(0.0);
			solve_info fi = slv(rhs, y);
			A.residual(rhs, y, r);
			const double rel = r.l2norm().get() / rhs.l2norm().get();
			const double want = name == std::string("krylov_cg.cfg") ? 1e-6 : 1e-10;
			std::printf("factory %s: %d iterations, %d restarts, true rel res %.3e\n", name,
			            fi.iters, fi.restarts, rel);
			fails += !fi.success() || rel > 1.01 * want;
		}
	}
	if (argc > 1) {
		// the reference's test input path: read, tocsr, device operator, CG
		auto coo = mat::io::matrix_market<double, int32_t>::read(argv[1]);
		auto csr = coo.tocsr();
		op::core<mat::device_csr> M(csr.device(ctx));
		const int m = (int)csr.rows();
		vec::device<double> mb(ctx, m), mx(ctx, m), mr(ctx, m);
		mb.set_scalar(1.0);
		cg::settings ms;
		ms.maxiter = 100;
		ms.rtol = 1e-12;
		ms.use_zero_guess = true;
		auto mslv = cg::solver(ms, vec::make_device_work<cg::nwork>(mb))(op::ref(M));
		solve_info mi = mslv(mb, mx);
		M.residual(mb, mx, mr);
		const double mrel = mr.l2norm().get() / mb.l2norm().get();
		std::printf("matrix market %s: %d rows, %d entries, cg %d iterations, true rel res %.3e\n",
		            argv[1], m, (int)csr.nnz(), mi.iters, mrel);
		fails += !mi.success() || mrel > 1e-11 || csr.nnz() != (int32_t)coo.nnz();
	}
	fsx_csr_destroy(Ah);
	fsx_ctx_destroy(ctx);
	std::printf(fails ? "FAIL\n" : "PASS\n");
	return fails ? 1 : 0;
}
