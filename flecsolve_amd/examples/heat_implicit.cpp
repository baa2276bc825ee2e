// Implicit 2D heat equation: the reference's heat example
// (examples/heat_equation/implicit.cc, heat.hh, implicit.cfg) on one GPU.
//
//   u_t = alpha Laplacian(u) on (0,10)^2, u = 0 on the boundary,
//   u(0) = 50 on [4,6]^2 and 0 elsewhere (examples/heat_equation/README.md),
//
// on an n x n grid of interior vertices (spacing 10/(n+1); the reference's
// mesh extents count the boundary vertices too, so its "N N" run is n = N-2
// here).  The time integrator and the stage solver come from the config file
// exactly as in the reference: read_config(file,
// bdf::options("time-integrator"), krylov_factory::options("linear-solver")).
//
// Two ways to run the stage solves:
//   generic (default): operator_adapter over the device CSR of alpha
//     Laplacian -- each product an SpMV plus an axpy -- and the solver
//     krylov_factory builds, the reference's structure with one host round
//     trip per reduction;
//   --fused: csr_adapter, which keeps I - gamma F as a device CSR (rewritten
//     by one kernel when gamma changes), solved by the fused graph-captured
//     solver of the configured type (cg, gmres or bicgstab).
//
// --explicit instead runs the reference's explicit example
// (examples/heat_equation/explicit.cc with its explicit.cfg): RK23 with the
// SpMV as f, no linear solves.
//
// usage: heat_implicit <cfg> <n> [--alpha A] [--fused | --explicit] [--out file]
//                      [--quiet] [--repeat R]
// Prints one line per accepted step (unless --quiet) and a summary line
//   "heat n N steps S rejects R solves K iters I time T wall W"
// with W the seconds spent in the time loop; --repeat runs the whole
// integration R times from u(0) (one summary line each; the first includes
// the one-time graph captures).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <string>
#include <vector>

#include "flecsolve/matrices/device_csr.hh"
#include "flecsolve/solvers/factory.hh"
#include "flecsolve/solvers/fused.hh"
#include "flecsolve/time-integrators/bdf.hh"
#include "flecsolve/time-integrators/operator_adapter.hh"
#include "flecsolve/time-integrators/rk23.hh"
#include "flecsolve/vectors/device.hh"
#include "fsx_problems.h"

using namespace flecsolve;
namespace ti = flecsolve::time_integrator;
using dvec = vec::device<double>;
using dcore = dvec::base; // the type operators are applied to

namespace {

struct heat_problem {
	std::vector<int32_t> rp, ci;
	std::vector<double> f, u0;
};

// f(u) = alpha Laplacian(u) as a CSR on the interior vertices
heat_problem make_heat(int n, double alpha) {
	heat_problem p;
	p.rp.resize((std::size_t)n * n + 1);
	p.ci.resize(fsx_gen_poisson2d_nnz(n));
	p.f.resize(p.ci.size());
	fsx_gen_poisson2d(n, p.rp.data(), p.ci.data(), p.f.data());
	const double h = 10.0 / (n + 1);
	const double c = alpha / (h * h);
	for (auto & e : p.f) e = -c * e;
	p.u0.resize((std::size_t)n * n);
	for (int j = 0; j < n; ++j)
		for (int i = 0; i < n; ++i) {
			const double x = (i + 1) * h, y = (j + 1) * h;
			p.u0[(std::size_t)j * n + i] = (x >= 4 && x <= 6 && y >= 4 && y <= 6) ? 50.0 : 0.0;
		}
	return p;
}

// stage solver with its solve and iteration counts
struct counted_solver : op::base<> {
	template<class B, class X>
	solve_info apply(const B & b, X & x) const {
		auto info = fn(b, x);
		iters += info.iters;
		++solves;
		return info;
	}
	std::function<solve_info(const dcore &, dcore &)> fn;
	mutable long iters = 0, solves = 0;
};

// download u to `out` (n*n doubles) when a file is named
void write_solution(const dvec & u, const char * out, fsx_ctx * ctx) {
	if (!out) return;
	std::vector<double> h(u.data.size());
	fsx_check(fsx_vec_download(ctx, u.handle(), h.data(), h.size()), "download");
	std::ofstream(out, std::ios::binary)
		.write(reinterpret_cast<const char *>(h.data()), (std::streamsize)(h.size() * 8));
}

// work vectors copy as aliases: every repeat reuses the same device buffers
// (and so the solver graphs captured for them)
template<class F, class S, class W>
int run(const ti::bdf::settings & ts, F op, S solver, const W & work, dvec & u, dvec & unew,
        int n, bool quiet, const char * out, fsx_ctx * ctx) {
	solver.get().iters = 0;
	solver.get().solves = 0;
	ti::bdf::integrator t(ti::bdf::parameters(ts, op, W(work), solver));
	// the loop of examples/heat_equation/implicit.cc:41-55
	const auto t0 = std::chrono::steady_clock::now();
	auto dt = t.get_current_dt();
	bool first_step = true;
	while (t.get_current_time() < t.get_final_time()) {
		t.advance(dt, first_step, u, unew);
		const bool good = t.check_solution();
		if (good) {
			if (!quiet)
				std::printf("Step %d advanced %gs to time %gs\n", t.get_current_step(), dt,
				            t.get_current_time());
			t.update();
			std::swap(u, unew);
			first_step = false;
		}
		dt = t.get_next_dt(good);
	}
	fsx_check(fsx_sync(ctx), "sync");
	const double wall =
		std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	write_solution(u, out, ctx);
	std::printf("heat n %d steps %d rejects %d solves %ld iters %ld time %.17g wall %.6f\n", n,
	            t.get_current_step(), t.num_step_rejects(), solver.get().solves,
	            solver.get().iters, t.get_current_time(), wall);
	return 0;
}

// the explicit example (examples/heat_equation/explicit.cc:18-48): RK23
// with the [time-integrator] section of its explicit.cfg, f(u) = F u
int run_explicit(const char * cfg, const heat_problem & p, int n, bool quiet,
                 const char * out, fsx_ctx * ctx, const mat::device_csr & Fm) {
	op::core<mat::device_csr> F(Fm);
	dvec u(ctx, p.u0.size()), unew(ctx, p.u0.size());
	fsx_check(fsx_vec_upload(ctx, u.handle(), p.u0.data(), p.u0.size()), "upload");
	ti::rk23::integrator t(ti::rk23::parameters(
		read_config(cfg, ti::rk23::options("time-integrator")), op::ref(F),
		ti::rk23::make_work(u)));
	const auto t0 = std::chrono::steady_clock::now();
	auto dt = t.get_current_dt();
	while (t.get_current_time() < t.get_final_time()) {
		t.advance(dt, u, unew);
		const bool good = t.check_solution();
		if (good) {
			if (!quiet)
				std::printf("Step %d advanced %gs to time %gs\n", t.get_current_step(), dt,
				            t.get_current_time());
			t.update();
			std::swap(u, unew);
		}
		dt = t.get_next_dt(good);
	}
	fsx_check(fsx_sync(ctx), "sync");
	const double wall =
		std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	write_solution(u, out, ctx);
	std::printf("heat n %d steps %d rejects %d solves 0 iters 0 time %.17g wall %.6f\n", n,
	            t.get_current_step(), t.num_step_rejects(), t.get_current_time(), wall);
	return 0;
}

}

int main(int argc, char ** argv) {
	if (argc < 3) {
		std::fprintf(stderr,
		             "usage: heat_implicit <cfg> <n> [--alpha A] [--fused | --explicit] "
		             "[--out file] [--quiet] [--repeat R]\n");
		return 2;
	}
	const char * cfg = argv[1];
	const int n = std::atoi(argv[2]);
	double alpha = 1.0;
	bool fused = false, quiet = false, expl = false;
	const char * out = nullptr;
	int repeat = 1;
	for (int i = 3; i < argc; ++i) {
		if (!std::strcmp(argv[i], "--alpha") && i + 1 < argc) alpha = std::atof(argv[++i]);
		else if (!std::strcmp(argv[i], "--fused")) fused = true;
		else if (!std::strcmp(argv[i], "--explicit")) expl = true;
		else if (!std::strcmp(argv[i], "--quiet")) quiet = true;
		else if (!std::strcmp(argv[i], "--out") && i + 1 < argc) out = argv[++i];
		else if (!std::strcmp(argv[i], "--repeat") && i + 1 < argc) repeat = std::atoi(argv[++i]);
		else {
			std::fprintf(stderr, "unknown argument %s\n", argv[i]);
			return 2;
		}
	}
	try {
		if (n < 1) throw std::invalid_argument("n must be positive");
		const auto p = make_heat(n, alpha);
		fsx_ctx * ctx;
		fsx_check(fsx_ctx_create(0, &ctx), "ctx");
		fsx_csr * Fh;
		fsx_check(fsx_csr_create(ctx, (int64_t)n * n, (int64_t)n * n, (int64_t)p.ci.size(),
		                         FSX_I32, p.rp.data(), p.ci.data(), FSX_F64, p.f.data(), &Fh),
		          "csr");
		const auto Fm = mat::device_csr::adopt(ctx, Fh);
		if (expl) return run_explicit(cfg, p, n, quiet, out, ctx, Fm);
		auto [ts, ss] = read_config(cfg, ti::bdf::options("time-integrator"),
		                            krylov_factory::options("linear-solver"));
		dvec u(ctx, p.u0.size()), unew(ctx, p.u0.size());
		const auto work = ti::bdf::make_work(u);
		auto solver = op::make_shared<counted_solver>();
		int rc = 0;
		if (repeat < 1) throw std::invalid_argument("--repeat must be positive");
		if (fused) {
			auto F = op::make_shared<ti::csr_adapter>(Fm);
			const auto A = F.get().matrix();
			switch (ss.target_id.value()) {
			case krylov_target::gmres: {
				gmres::fused g(std::get<gmres::settings>(ss.target_settings), A);
				solver.get().fn = [g](const dcore & b, dcore & x) { return g.apply(b, x); };
				break;
			}
			case krylov_target::bicgstab: {
				bicgstab::fused g(std::get<bicgstab::settings>(ss.target_settings), A);
				solver.get().fn = [g](const dcore & b, dcore & x) { return g.apply(b, x); };
				break;
			}
			default: {
				cg::fused g(std::get<cg::settings>(ss.target_settings), A);
				solver.get().fn = [g](const dcore & b, dcore & x) { return g.apply(b, x); };
			}
			}
			for (int r = 0; r < repeat; ++r) {
				fsx_check(fsx_vec_upload(ctx, u.handle(), p.u0.data(), p.u0.size()), "upload");
				rc = run(ts, F, solver, work, u, unew, n, quiet, out, ctx);
			}
		}
		else {
			auto F = op::make_shared<ti::operator_adapter<mat::device_csr>>(Fm);
			auto inner = krylov_factory::make_shared(ss, u, F);
			solver.get().fn = [inner](const dcore & b, dcore & x) { return inner.get().apply(b, x); };
			for (int r = 0; r < repeat; ++r) {
				fsx_check(fsx_vec_upload(ctx, u.handle(), p.u0.data(), p.u0.size()), "upload");
				rc = run(ts, F, solver, work, u, unew, n, quiet, out, ctx);
			}
		}
		return rc;
	}
	catch (const std::exception & e) {
		std::fprintf(stderr, "error: %s\n", e.what());
		return 1;
	}
}
