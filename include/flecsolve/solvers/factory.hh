// flecsolve/solvers/factory.hh -- run-time choice of Krylov solver.
//
// Drop-in for the reference's krylov_factory (flecsolve/solvers/factory.hh:
// 26-89, over operators/factory.hh:26-220): a config file names the solver
// with `<label>.type` (cg, gmres, bicgstab) and gives its options under
// `<label>.options` (the keys of cg::options, gmres::options,
// bicgstab::options); read_config(file, krylov_factory::options(label))
// returns krylov_factory::settings, and krylov_factory::make(settings,
// exemplar, A[, P[, diagnostic]]) builds that solver with work vectors
// shaped like the exemplar and returns an operator whose apply(b, x) runs it
// and returns its solve_info; make_shared(...) returns it behind an
// op::handle, as the time integrators take their solver.  The reference's fourth
// target, nka (a nonlinear Krylov accelerator), is not part of the linear
// solve path built here.
#ifndef FLECSOLVE_AMD_SOLVERS_FACTORY_HH
#define FLECSOLVE_AMD_SOLVERS_FACTORY_HH

#include <istream>
#include <optional>
#include <string>
#include <variant>

#include "flecsolve/solvers/bicgstab.hh"
#include "flecsolve/solvers/cg.hh"
#include "flecsolve/solvers/gmres.hh"
#include "flecsolve/operators/handle.hh"
#include "flecsolve/util/config.hh"
#include "flecsolve/vectors/util.hh"

namespace flecsolve {

enum class krylov_target { cg, gmres, bicgstab };

inline std::istream & operator>>(std::istream & in, krylov_target & t) {
	std::string tok;
	in >> tok;
	if (tok == "cg") t = krylov_target::cg;
	else if (tok == "gmres") t = krylov_target::gmres;
	else if (tok == "bicgstab") t = krylov_target::bicgstab;
	else in.setstate(std::ios_base::failbit);
	return in;
}

template<krylov_target T>
struct krylov_registry;

template<>
struct krylov_registry<krylov_target::cg> {
	using settings = cg::settings;
	using options = cg::options;
	template<class V, class... Args>
	static auto make(const settings & s, const V & v, Args &&... args) {
		return cg::solver(s, vec::make_work<cg::nwork>(v))(std::forward<Args>(args)...);
	}
};

template<>
struct krylov_registry<krylov_target::gmres> {
	using settings = gmres::settings;
	using options = gmres::options;
	template<class V, class... Args>
	static auto make(const settings & s, const V & v, Args &&... args) {
		return gmres::solver(s, vec::make_work<gmres::nwork>(v))(
			std::forward<Args>(args)...);
	}
};

template<>
struct krylov_registry<krylov_target::bicgstab> {
	using settings = bicgstab::settings;
	using options = bicgstab::options;
	template<class V, class... Args>
	static auto make(const settings & s, const V & v, Args &&... args) {
		return bicgstab::solver(s, vec::make_work<bicgstab::nwork>(v))(
			std::forward<Args>(args)...);
	}
};

// The made operator: apply visits the solver chosen at run time
// (operators/factory.hh:28-53).
template<class Variant>
struct factory_prod : op::base<> {
	explicit factory_prod(Variant v) : var(std::move(v)) {}
	template<class D, class R>
	decltype(auto) apply(const D & x, R & y) const {
		return std::visit([&](const auto & p) { return p.apply(x, y); }, var);
	}
	template<class D, class R>
	decltype(auto) operator()(const D & x, R & y) const {
		return apply(x, y);
	}
	Variant var;
};

struct krylov_factory {
	using target = krylov_target;
	using settings_types =
		std::variant<cg::settings, gmres::settings, bicgstab::settings>;

	struct settings {
		std::optional<target> target_id;
		settings_types target_settings;
	};

	static settings_types make_settings(target t) {
		switch (t) {
		case target::gmres: return gmres::settings{};
		case target::bicgstab: return bicgstab::settings{};
		default: return cg::settings{};
		}
	}

	static optdesc make_options(const std::string & pre, settings_types & s) {
		return std::visit(
			[&](auto & st) -> optdesc {
				using S = std::decay_t<decltype(st)>;
				if constexpr (std::is_same_v<S, gmres::settings>)
					return gmres::options(pre.c_str())(st);
				else if constexpr (std::is_same_v<S, bicgstab::settings>)
					return bicgstab::options(pre.c_str())(st);
				else
					return cg::options(pre.c_str())(st);
			},
			s);
	}

	// `<label>.type` (required) and, once it is known, the chosen solver's
	// options under `<label>.options` (operators/factory.hh:93-118)
	struct options : with_label {
		using settings_type = settings;
		explicit options(const char * pre) : with_label(pre) {}
		optdesc operator()(settings & s) {
			optdesc d;
			d.required<target>(label("type"), &parsed,
			                   [&s](const target & t) { s.target_id.emplace(t); });
			if (s.target_id.has_value()) {
				s.target_settings = make_settings(s.target_id.value());
				d.add(make_options(label("options"), s.target_settings));
			}
			return d;
		}
		target parsed = target::cg;
	};

	// the made operator behind a shared handle (operators/factory.hh:148-153)
	template<class V, class... Args>
	static auto make_shared(const settings & s, const V & exemplar, Args &&... args) {
		auto p = make(s, exemplar, std::forward<Args>(args)...);
		using prod = decltype(p);
		return op::handle<op::core<prod>>{std::make_shared<op::core<prod>>(std::move(p))};
	}

	// operators/factory.hh:141-177
	template<class V, class... Args>
	static auto make(const settings & s, const V & exemplar, Args &&... args) {
		using var_t = std::variant<
			decltype(krylov_registry<target::cg>::make(cg::settings{}, exemplar, args...)),
			decltype(krylov_registry<target::gmres>::make(gmres::settings{}, exemplar, args...)),
			decltype(krylov_registry<target::bicgstab>::make(bicgstab::settings{}, exemplar,
			                                                  args...))>;
		if (!s.target_id) throw std::invalid_argument("krylov_factory: no solver type");
		switch (s.target_id.value()) {
		case target::gmres:
			return factory_prod<var_t>{var_t{std::in_place_index<1>,
			                                 krylov_registry<target::gmres>::make(
				                                 std::get<gmres::settings>(s.target_settings),
				                                 exemplar, args...)}};
		case target::bicgstab:
			return factory_prod<var_t>{var_t{std::in_place_index<2>,
			                                 krylov_registry<target::bicgstab>::make(
				                                 std::get<bicgstab::settings>(s.target_settings),
				                                 exemplar, args...)}};
		default:
			return factory_prod<var_t>{var_t{std::in_place_index<0>,
			                                 krylov_registry<target::cg>::make(
				                                 std::get<cg::settings>(s.target_settings),
				                                 exemplar, args...)}};
		}
	}
};

}

#endif
