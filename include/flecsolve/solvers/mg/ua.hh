// flecsolve/solvers/mg/ua.hh -- unsmoothed-aggregation AMG as a device
// operator.
//
// Drop-in for the reference's mg::ua::solver / bound_solver
// (flecsolve/solvers/mg/ua.hh:76-212, coarsen.hh:598-605, cycle.hh:7-12):
// same settings structs and cycle names; apply(b, x) sets x = 0 and runs up
// to maxiter cycles until |b - A x| <= atol.  The hierarchy is built by
// fsx_amg_create (pairwise aggregation and Galerkin coarse operators with the
// reference's ordering, dense LU on the coarsest level) and every cycle runs
// on the device.  One colour: the redistribution settings have nothing to
// act on, and the BoomerAMG coarse solver (AMP) is not available.
#ifndef FLECSOLVE_AMD_SOLVERS_MG_UA_HH
#define FLECSOLVE_AMD_SOLVERS_MG_UA_HH

#include <cstddef>
#include <memory>
#include <stdexcept>

#include "fsx.h"
#include "flecsolve/matrices/device_csr.hh"
#include "flecsolve/operators/core.hh"

namespace flecsolve::mg {

enum class cycle_type { v, relaxed_w, relaxed_kappa, kappa_k };

namespace ua {

struct coarsen_settings {
	float beta = 0.25;
	std::size_t redist_coarsen_factor = 2;
	std::size_t min_local_coarse = 5;
	std::size_t pairwise_passes = 2;
	bool coarsen_to_serial = false;
	bool redistribute = true;
};

// option defaults of the reference's ua::solver::options (ua.cc)
struct solver_settings {
	std::size_t max_levels = 20;
	std::size_t min_coarse = 5;
	std::size_t maxiter = 10;
	float atol = 1e-6;
	coarsen_settings coarsening_settings{};
	float jacobi_weight = 0.6666666;
	std::size_t nrelax = 2;
	cycle_type cycle = cycle_type::v;
	bool boomer_cg = false;
	int kappa = 1;
	float ktol = 0.25;
};

inline fsx_amg_settings to_fsx(const solver_settings & s) {
	if (s.boomer_cg) throw std::invalid_argument("ua: the BoomerAMG coarse solver is not built");
	fsx_amg_settings o;
	o.max_levels = (int)s.max_levels;
	o.min_coarse = (int64_t)s.min_coarse;
	o.maxiter = (int)s.maxiter;
	o.atol = s.atol;
	o.beta = s.coarsening_settings.beta;
	o.pairwise_passes = (int)s.coarsening_settings.pairwise_passes;
	o.nrelax = (int)s.nrelax;
	o.cycle = (int)s.cycle;
	o.kappa = s.kappa;
	o.ktol = s.ktol;
	o.smoother = 0; // hybrid Gauss-Seidel, as create_smoother (ua.hh:184-197)
	o.jacobi_weight = s.jacobi_weight;
	return o;
}

struct bound_solver : op::base<> {
	bound_solver(mat::device_csr a, const solver_settings & s) : A(std::move(a)), settings(s) {
		const fsx_amg_settings fs = to_fsx(s);
		fsx_amg * m = nullptr;
		fsx_check(fsx_amg_create(A.context(), A.handle(), &fs, &m), "amg_create");
		M = std::shared_ptr<fsx_amg>(m, [](fsx_amg * p) { fsx_amg_destroy(p); });
	}
	// wrap an existing solver (not owned)
	bound_solver(mat::device_csr a, fsx_amg * m) : A(std::move(a)), M(m, [](fsx_amg *) {}) {}

	template<class D, class R>
	void apply(const D & b, R & x) const {
		fsx_check(fsx_amg_apply(A.context(), M.get(), b.data.get(), x.data.get(), nullptr,
		                        nullptr),
		          "amg_apply");
	}

	mat::device_csr A; // keeps the fine matrix alive
	solver_settings settings;
	std::shared_ptr<fsx_amg> M;
};

struct solver {
	using settings = solver_settings;
	explicit solver(const settings & s) : settings_(s) {}
	auto operator()(mat::device_csr A) const { return op::core<bound_solver>{A, settings_}; }
	settings settings_;
};

}
}

#endif
