// fsx_amg_setup.hpp -- host side of the UA-AMG setup (fsx_amg.hip): pairwise
// aggregation, the auxiliary and Galerkin coarse matrices, the dense LU of
// the coarsest level.  Plain C++ (no HIP), so the setup is also built and
// checked on the host alone.  The semantics restated here are listed at the
// top of fsx_amg.hip.
//
// Work that is independent per row (the candidate test, the strength graph,
// the coarse rows) runs on up to kSetupThreads host threads over contiguous
// row blocks, each row computed exactly as the serial loop would and the
// blocks concatenated in row order, so the results do not depend on the
// thread count.  The priority-queue walk that forms the aggregates is
// serial: its order is the algorithm.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <exception>
#include <limits>
#include <thread>
#include <vector>

#include "fsx_host_threads.hpp"

namespace fsx {

// ---- host-side structures ---------------------------------------------------

struct hcsr {
	int64_t n = 0;
	std::vector<int64_t> rp{0}, ci;
	std::vector<double> v;
};

constexpr int64_t kNoAgg = -1;

// Aggregates, flat: aggregate i holds idx[ptr[i] .. ptr[i+1]) in the order
// its members were added.
struct aggregates {
	std::vector<int64_t> ptr{0}, idx;
	struct range {
		const int64_t *b, *e;
		const int64_t * begin() const { return b; }
		const int64_t * end() const { return e; }
		size_t size() const { return (size_t)(e - b); }
	};
	size_t size() const { return ptr.size() - 1; }
	range operator[](size_t i) const {
		return {idx.data() + ptr[i], idx.data() + ptr[i + 1]};
	}
	void push(int64_t a) {
		idx.push_back(a);
		ptr.push_back((int64_t)idx.size());
	}
	void push(int64_t a, int64_t b) {
		idx.push_back(a);
		idx.push_back(b);
		ptr.push_back((int64_t)idx.size());
	}
};

// Candidates keyed by (priority, insertion sequence): the order of the
// reference's multimap.  After build() priorities only decrease, so each
// priority keeps a first-in-first-out array of its candidates and a moved or
// removed candidate's old entry is skipped when reached (it is stale exactly
// when the candidate's priority is no longer that array's): appends happen in
// insertion order, so each array's live entries are in (priority, insertion)
// order and pop() takes the first live entry of the lowest priority.
class prospect {
public:
	explicit prospect(int64_t n) : prio_((size_t)n, -1) {}
	bool contains(int64_t e) const { return prio_[(size_t)e] >= 0; }
	bool empty() const { return count_ == 0; }
	void add(int64_t e) { // before build(): priority 0
		prio_[(size_t)e] = 0;
		++count_;
	}
	void bump(int64_t e) { ++prio_[(size_t)e]; }
	// append every candidate in ascending id (the multimap built from the
	// id-ordered map)
	void build() {
		for (int64_t e = 0; e < (int64_t)prio_.size(); ++e)
			if (prio_[(size_t)e] >= 0) push_back(e);
	}
	int64_t pop() {
		for (;;) {
			auto & q = fifo_[(size_t)lowest_];
			size_t & h = head_[(size_t)lowest_];
			while (h < q.size() && prio_[(size_t)q[h]] != lowest_) ++h;
			if (h < q.size()) {
				const int64_t e = q[h++];
				prio_[(size_t)e] = -1;
				--count_;
				return e;
			}
			++lowest_;
		}
	}
	void remove(int64_t e) {
		prio_[(size_t)e] = -1;
		--count_;
	}
	void decrement(int64_t e) {
		if (!contains(e)) return;
		--prio_[(size_t)e];
		push_back(e);
	}

private:
	void push_back(int64_t e) {
		const int64_t p = prio_[(size_t)e];
		if (p >= (int64_t)fifo_.size()) {
			fifo_.resize((size_t)p + 1);
			head_.resize((size_t)p + 1, 0);
		}
		fifo_[(size_t)p].push_back(e);
		lowest_ = std::min(lowest_, p);
	}
	std::vector<int64_t> prio_;
	std::vector<std::vector<int64_t>> fifo_;
	std::vector<size_t> head_;
	int64_t lowest_ = 0, count_ = 0;
};

// strong-neighbour sets: sorted, no duplicates (std::set order)
struct soc_sets {
	std::vector<int64_t> ptr{0}, idx;
	bool has(int64_t r, int64_t c) const {
		return std::binary_search(idx.begin() + ptr[r], idx.begin() + ptr[r + 1], c);
	}
};

// Rows assembled per thread block, then concatenated in row order:
// per row, its entry count; per block, the entries.
template<class T>
struct row_blocks {
	std::vector<int64_t> cnt;              // entries per row
	std::vector<std::vector<T>> part;      // per block
	explicit row_blocks(int64_t n, unsigned nt) : cnt((size_t)n, 0), part(nt) {}
	// row pointer (size n + 1) and the concatenated entries
	void join(std::vector<int64_t> & ptr, std::vector<T> & out) const {
		const size_t n = cnt.size();
		ptr.assign(n + 1, 0);
		for (size_t r = 0; r < n; ++r) ptr[r + 1] = ptr[r] + cnt[r];
		out.clear();
		out.reserve((size_t)ptr[n]);
		for (const auto & p : part) out.insert(out.end(), p.begin(), p.end());
	}
};

inline aggregates pairwise_agg(float beta, const hcsr & A, bool checkdd) {
	const int64_t n = A.n;
	const unsigned nt = setup_threads(n);
	prospect unmarked(n);
	if (checkdd) {
		std::vector<char> cand((size_t)n, 0);
		for_row_blocks(n, nt, [&](unsigned, int64_t lo, int64_t hi) {
			for (int64_t r = lo; r < hi; ++r) {
				int offd_sum = 0;
				double diag_val = 0.;
				for (int64_t off = A.rp[r]; off < A.rp[r + 1]; ++off) {
					if (r != A.ci[off]) offd_sum += std::abs(A.v[off]);
					else diag_val = A.v[off];
				}
				cand[r] = !(diag_val > 5 * offd_sum);
			}
		});
		for (int64_t r = 0; r < n; ++r)
			if (cand[r]) unmarked.add(r);
	}
	else
		for (int64_t r = 0; r < n; ++r) unmarked.add(r);
	soc_sets soc;
	{
		const int64_t nnz = (int64_t)A.v.size();
		row_blocks<int64_t> rb(n, nt);
		for_row_blocks(n, nt, [&](unsigned t, int64_t lo, int64_t hi) {
			auto & out = rb.part[t];
			std::vector<int64_t> tmp;
			for (int64_t r = lo; r < hi; ++r) {
				// smallest stored value of the row; an empty row reads the next
				// stored value, past the last the largest double (min_element's
				// end)
				double maxc;
				if (A.rp[r] < A.rp[r + 1])
					maxc = *std::min_element(A.v.begin() + A.rp[r], A.v.begin() + A.rp[r + 1]);
				else
					maxc = A.rp[r + 1] < nnz ? A.v[A.rp[r + 1]]
					                         : std::numeric_limits<double>::max();
				tmp.clear();
				for (int64_t off = A.rp[r]; off < A.rp[r + 1]; ++off) {
					const int64_t c = A.ci[off];
					if (r != c && (!checkdd || unmarked.contains(c)) && c < n &&
					    A.v[off] < (beta * maxc))
						tmp.push_back(c);
				}
				std::sort(tmp.begin(), tmp.end());
				tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
				out.insert(out.end(), tmp.begin(), tmp.end());
				rb.cnt[r] = (int64_t)tmp.size();
			}
		});
		rb.join(soc.ptr, soc.idx);
	}
	for (int64_t r = 0; r < n; ++r)
		for (int64_t q = soc.ptr[r]; q < soc.ptr[r + 1]; ++q) unmarked.bump(soc.idx[q]);
	unmarked.build();

	aggregates agg;
	agg.ptr.reserve((size_t)n / 2 + 2);
	agg.idx.reserve((size_t)n);
	auto update_priorities = [&](int64_t k) {
		for (int64_t q = soc.ptr[k]; q < soc.ptr[k + 1]; ++q) unmarked.decrement(soc.idx[q]);
	};
	while (!unmarked.empty()) {
		const int64_t sel = unmarked.pop();
		double best = std::numeric_limits<double>::max();
		int64_t pair = -1;
		for (int64_t off = A.rp[sel]; off < A.rp[sel + 1]; ++off) {
			const int64_t c = A.ci[off];
			if (unmarked.contains(c) && A.v[off] < best) {
				best = A.v[off];
				pair = c;
			}
		}
		if (pair >= 0 && soc.has(sel, pair)) {
			unmarked.remove(pair);
			update_priorities(pair);
			agg.push(sel, pair);
		}
		else
			agg.push(sel);
		update_priorities(sel);
	}
	return agg;
}

// A coarse row from its (column, value) pairs in traversal order: the
// columns ascending, each column's values summed in traversal order starting
// from 0.0 (the reference's std::map accumulation).  The row's distinct
// columns are kept sorted while the pairs arrive.
struct row_accumulator {
	std::vector<int64_t> col;
	std::vector<double> sum;
	void clear() {
		col.clear();
		sum.clear();
	}
	void add(int64_t c, double v) {
		auto it = std::lower_bound(col.begin(), col.end(), c);
		const size_t p = (size_t)(it - col.begin());
		if (it != col.end() && *it == c) {
			sum[p] += v;
			return;
		}
		col.insert(it, c);
		sum.insert(sum.begin() + (std::ptrdiff_t)p, 0.0 + v);
	}
};

struct coarse_entry {
	int64_t c;
	double v;
};

// The coarse matrix of the aggregates: row i sums the rows of aggregate i,
// column ids mapped through col_of (entries mapped to kNoAgg dropped).
template<class ColOf>
hcsr coarse_rows(const hcsr & A, const aggregates & agg, ColOf && col_of) {
	hcsr C;
	C.n = (int64_t)agg.size();
	const unsigned nt = setup_threads(C.n);
	row_blocks<coarse_entry> rb(C.n, nt);
	for_row_blocks(C.n, nt, [&](unsigned t, int64_t lo, int64_t hi) {
		auto & out = rb.part[t];
		row_accumulator acc;
		for (int64_t rc = lo; rc < hi; ++rc) {
			acc.clear();
			for (int64_t r : agg[(size_t)rc])
				for (int64_t off = A.rp[r]; off < A.rp[r + 1]; ++off) {
					const int64_t a = col_of(A.ci[off]);
					if (a != kNoAgg) acc.add(a, A.v[off]);
				}
			for (size_t j = 0; j < acc.col.size(); ++j) out.push_back({acc.col[j], acc.sum[j]});
			rb.cnt[rc] = (int64_t)acc.col.size();
		}
	});
	std::vector<coarse_entry> all;
	rb.join(C.rp, all);
	C.ci.resize(all.size());
	C.v.resize(all.size());
	for (size_t e = 0; e < all.size(); ++e) {
		C.ci[e] = all[e].c;
		C.v[e] = all[e].v;
	}
	return C;
}

// fine row -> aggregate id (kNoAgg outside every aggregate)
inline std::vector<int64_t> aggregate_of(int64_t n, const aggregates & agg) {
	std::vector<int64_t> aggt((size_t)n, kNoAgg);
	for (size_t a = 0; a < agg.size(); ++a)
		for (int64_t f : agg[a]) aggt[(size_t)f] = (int64_t)a;
	return aggt;
}

// create_aux: aggregate ids through a map that answers 0 for rows outside
// every aggregate (coarsen.hh:280-303)
inline hcsr create_aux(const hcsr & A, const aggregates & agg) {
	std::vector<int64_t> aggt = aggregate_of(A.n, agg);
	for (auto & a : aggt)
		if (a == kNoAgg) a = 0;
	return coarse_rows(A, agg, [&](int64_t c) { return aggt[(size_t)c]; });
}

inline aggregates agg_union(const aggregates & a1, const aggregates & a2) {
	aggregates a;
	a.ptr.resize(a2.size() + 1);
	a.ptr[0] = 0;
	for (size_t i = 0; i < a2.size(); ++i) {
		int64_t s = 0;
		for (int64_t e2 : a2[i]) s += (int64_t)a1[(size_t)e2].size();
		a.ptr[i + 1] = a.ptr[i] + s;
	}
	a.idx.resize((size_t)a.ptr.back());
	int64_t k = 0;
	for (size_t i = 0; i < a2.size(); ++i)
		for (int64_t e2 : a2[i])
			for (int64_t e1 : a1[(size_t)e2]) a.idx[(size_t)k++] = e1;
	return a;
}

inline aggregates multi_pairwise_agg(float beta, const hcsr & A, int passes) {
	aggregates agg1 = pairwise_agg(beta, A, true);
	hcsr aux = create_aux(A, agg1);
	aggregates agg2 = pairwise_agg(beta, aux, false);
	if (passes == 2) return agg_union(agg1, agg2);
	hcsr aux1 = create_aux(aux, agg2);
	aggregates agg3 = pairwise_agg(beta, aux1, false);
	return agg_union(agg_union(agg1, agg2), agg3);
}

// coarsen_with_aggregates (coarsen.hh:455-488): aggt = kNoAgg outside
// every aggregate, those columns dropped
inline hcsr galerkin(const hcsr & A, const aggregates & agg, const std::vector<int64_t> & aggt) {
	return coarse_rows(A, agg, [&](int64_t c) { return aggt[(size_t)c]; });
}

// bound_solver::setup's coarsening (ua.hh:110-139): from the fine matrix H,
// aggregate and form the Galerkin matrix level by level until a level has
// fewer than min_coarse rows or max_levels levels exist; level(i, F, agg,
// aggt, C) sees each step (fine matrix, aggregates, fine row -> aggregate,
// coarse matrix).  Returns the coarsest matrix.
template<class F>
hcsr coarsen_levels(hcsr H, float beta, int passes, int max_levels, int64_t min_coarse,
                    F && level) {
	for (int i = 0; i + 1 < max_levels; ++i) {
		aggregates agg = multi_pairwise_agg(beta, H, passes);
		std::vector<int64_t> aggt = aggregate_of(H.n, agg);
		hcsr C = galerkin(H, agg, aggt);
		level(i, (const hcsr &)H, (const aggregates &)agg, (const std::vector<int64_t> &)aggt,
		      (const hcsr &)C);
		H = std::move(C);
		if (H.n < min_coarse) break;
	}
	return H;
}

// Dense LU with partial pivoting (row of the largest |a_ik|, the first on
// ties), row-major, in place; then the two triangular solves.
struct dense_lu {
	int64_t n = 0;
	std::vector<double> a;
	std::vector<int64_t> piv;
	void factor(const hcsr & A) {
		n = A.n;
		a.assign((size_t)(n * n), 0.0);
		for (int64_t r = 0; r < n; ++r)
			for (int64_t off = A.rp[r]; off < A.rp[r + 1]; ++off)
				a[(size_t)(r * n + A.ci[off])] = A.v[off];
		piv.assign((size_t)n, 0);
		for (int64_t k = 0; k < n; ++k) {
			int64_t p = k;
			double m = std::abs(a[(size_t)(k * n + k)]);
			for (int64_t i = k + 1; i < n; ++i)
				if (std::abs(a[(size_t)(i * n + k)]) > m) {
					m = std::abs(a[(size_t)(i * n + k)]);
					p = i;
				}
			piv[(size_t)k] = p;
			if (p != k)
				for (int64_t j = 0; j < n; ++j) std::swap(a[(size_t)(k * n + j)], a[(size_t)(p * n + j)]);
			const double d = a[(size_t)(k * n + k)];
			if (d == 0.0) continue;
			for (int64_t i = k + 1; i < n; ++i) {
				const double l = a[(size_t)(i * n + k)] / d;
				a[(size_t)(i * n + k)] = l;
				for (int64_t j = k + 1; j < n; ++j)
					a[(size_t)(i * n + j)] = a[(size_t)(i * n + j)] - l * a[(size_t)(k * n + j)];
			}
		}
	}
	void solve(const double * b, double * x) const {
		std::vector<double> y(b, b + n);
		for (int64_t k = 0; k < n; ++k) std::swap(y[(size_t)k], y[(size_t)piv[(size_t)k]]);
		for (int64_t i = 0; i < n; ++i)
			for (int64_t j = 0; j < i; ++j) y[(size_t)i] = y[(size_t)i] - a[(size_t)(i * n + j)] * y[(size_t)j];
		for (int64_t i = n - 1; i >= 0; --i) {
			for (int64_t j = i + 1; j < n; ++j) y[(size_t)i] = y[(size_t)i] - a[(size_t)(i * n + j)] * y[(size_t)j];
			y[(size_t)i] = y[(size_t)i] / a[(size_t)(i * n + i)];
		}
		std::copy(y.begin(), y.end(), x);
	}
};

} // namespace fsx
