// fsx_host_threads.hpp -- host threads for the library's setup work (the
// UA-AMG coarsening, the Gauss-Seidel level copies): contiguous row blocks
// or items taken in turn, each computed exactly as the serial loop would,
// so no result depends on the thread count.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <exception>
#include <thread>
#include <vector>

namespace fsx {

// Host threads for setup work over n rows: the hardware's, at most
// kSetupThreads, at most FSX_SETUP_THREADS when that is set, and at most one
// per 4096 rows.
constexpr unsigned kSetupThreads = 16;

inline unsigned setup_threads(int64_t n) {
	unsigned t = std::max(1u, std::min(kSetupThreads, std::thread::hardware_concurrency()));
	if (const char * e = std::getenv("FSX_SETUP_THREADS")) {
		const long v = std::strtol(e, nullptr, 10);
		if (v >= 1) t = std::min<unsigned>(t, (unsigned)v);
	}
	// at least 4096 rows per thread
	return (unsigned)std::max<int64_t>(1, std::min<int64_t>(t, n / 4096));
}

// body(t) on threads t = 0 .. nt-1, all joined before returning (also when
// starting a thread fails: the ones started are joined, then the error is
// rethrown).  body must not throw.
template<class B>
void run_threads(unsigned nt, B && body) {
	std::vector<std::thread> th;
	th.reserve(nt);
	try {
		for (unsigned t = 0; t < nt; ++t) th.emplace_back([&body, t] { body(t); });
	}
	catch (...) {
		for (auto & x : th) x.join();
		throw;
	}
	for (auto & x : th) x.join();
}

// fn(t, lo, hi) on rows [lo, hi): contiguous blocks, block t on thread t.
// An exception in a block is rethrown on the calling thread after the join.
template<class F>
void for_row_blocks(int64_t n, unsigned nt, F && fn) {
	if (nt <= 1) {
		fn(0u, (int64_t)0, n);
		return;
	}
	std::vector<std::exception_ptr> err(nt);
	run_threads(nt, [&](unsigned t) {
		try {
			fn(t, n * t / nt, n * (t + 1) / nt);
		}
		catch (...) {
			err[t] = std::current_exception();
		}
	});
	for (auto & e : err)
		if (e) std::rethrow_exception(e);
}

// fn(i) for every i in [0, count), the items taken in turn by nt threads.
// An exception is rethrown on the calling thread after the join.
template<class F>
void for_items(size_t count, unsigned nt, F && fn) {
	nt = (unsigned)std::min<size_t>(nt, count);
	if (nt <= 1) {
		for (size_t i = 0; i < count; ++i) fn(i);
		return;
	}
	std::atomic<size_t> next{0};
	std::vector<std::exception_ptr> err(nt);
	run_threads(nt, [&](unsigned t) {
		try {
			for (size_t i; (i = next.fetch_add(1)) < count;) fn(i);
		}
		catch (...) {
			err[t] = std::current_exception();
		}
	});
	for (auto & e : err)
		if (e) std::rethrow_exception(e);
}

} // namespace fsx
