// flecsolve/solvers/mg/gs.hh -- hybrid Gauss-Seidel relaxation on the device.
//
// Drop-in for the reference's op::hybrid_gs and mg::hybrid_gs::solver
// (flecsolve/solvers/mg/gs.hh:27-141): same settings (nrelax, sweep) and
// enums, apply(b, x) relaxes x in place.  The reference runs the relax task
// per colour over its parcsr diag/offd blocks; here the sweep is the
// libfsx level-scheduled kernel chain (fsx_gs_*), bit-identical to the
// sequential sweep, with `colours` row blocks (FleCSI equal_map partition)
// standing for the reference's ranks.
#ifndef FLECSOLVE_AMD_SOLVERS_MG_GS_HH
#define FLECSOLVE_AMD_SOLVERS_MG_GS_HH

#include <cstddef>
#include <memory>

#include "fsx.h"
#include "flecsolve/matrices/device_csr.hh"
#include "flecsolve/operators/core.hh"

namespace flecsolve {

namespace mg {
enum class relax_sweep { forward, backward, symmetric };
enum class relax_dir { forward, backward };

namespace hybrid_gs {
struct settings {
	std::size_t nrelax;
	relax_sweep sweep;
};
}
}

namespace op {
struct hybrid_gs : base<> {
	using settings_type = mg::hybrid_gs::settings;

	hybrid_gs(mat::device_csr a, const settings_type & s, int colours = 1)
		: A(std::move(a)), settings(s) {
		fsx_gs * g = nullptr;
		fsx_check(fsx_gs_create(A.context(), A.handle(), colours, &g), "gs_create");
		G = std::shared_ptr<fsx_gs>(g, [](fsx_gs * p) { fsx_gs_destroy(p); });
	}

	template<class D, class R>
	void apply(const D & b, R & x) const {
		const int sweep = settings.sweep == mg::relax_sweep::forward    ? 0
		                  : settings.sweep == mg::relax_sweep::backward ? 1
		                                                                : 2;
		fsx_check(fsx_gs_relax(A.context(), G.get(), sweep, (int)settings.nrelax,
		                       b.data.get(), x.data.get()),
		          "gs_relax");
	}

	mat::device_csr A; // keeps the matrix alive for the smoother
	settings_type settings;
	std::shared_ptr<fsx_gs> G;
};
}

namespace mg::hybrid_gs {
struct solver {
	auto operator()(mat::device_csr A) const { return op::core<op::hybrid_gs>{A, settings_}; }
	settings settings_;
};
}

}

#endif
