// flecsolve/operators/handle.hh -- shared or borrowed operator handles.
// Interface of the reference's flecsolve/operators/handle.hh:10-56
// (handle<T> over variant<shared_ptr, reference_wrapper>, ref, cref,
// make_shared).
#ifndef FLECSOLVE_AMD_OPERATORS_HANDLE_HH
#define FLECSOLVE_AMD_OPERATORS_HANDLE_HH

#include <functional>
#include <memory>
#include <variant>

#include "flecsolve/operators/core.hh"

namespace flecsolve::op {

template<class T>
struct handle {
	using type = T;
	std::variant<std::shared_ptr<T>, std::reference_wrapper<T>> store;

	T & get() const {
		if (auto * sp = std::get_if<std::shared_ptr<T>>(&store)) return **sp;
		return std::get<std::reference_wrapper<T>>(store).get();
	}
	operator T &() const { return get(); }
	template<class D, class R>
	decltype(auto) operator()(const D & x, R & y) const { return get()(x, y); }
};

template<class T>
auto ref(T & o) { return handle<T>{std::ref(o)}; }
template<class T>
auto cref(const T & o) { return handle<const T>{std::cref(o)}; }

template<class P, class... A>
auto make_shared(A &&... a) {
	return handle<core<P>>{std::make_shared<core<P>>(std::forward<A>(a)...)};
}
template<class P>
auto make_shared(core<P> && o) {
	return handle<core<P>>{std::make_shared<core<P>>(std::move(o))};
}

}

#endif
