// fsx_cg.hip -- fused CG driver with device-resident scalars, HIP-graph
// captured iterations, and FP32-inner / FP64-outer iterative refinement.
//
// Algorithm: op::cg::apply (flecsolve/solvers/cg.hh:43-139), unchanged in
// its iteration sequence, tolerance semantics and solve_info.  What changes
// is where the scalars live.  The reference forces a host round trip at every
// .get() (three per iteration: w.p, |r|, r.z -- cg.hh:98, 111, 127).  Here
// one iteration is three streaming kernels, each dot product finishing in a
// one-block kernel on the device:
//
//   K1  w = A p, partials of w.p           (cg.hh:95-98)    SpMV + dot
//   F1  alpha = rho / w.p                  (cg.hh:98-106)
//   K2  r -= a w; z = D^-1 r; partials of r.r and r.z
//                                          (cg.hh:109-127)  one HBM pass
//   F2  |r|, tolerance test, rho, beta     (cg.hh:111-129)
//   K3  x += a p; p = beta p + z           (cg.hh:108, 130)
//
// x += alpha p (cg.hh:108) is deferred from K2 to K3, which reads p anyway:
// the iteration moves 80 instead of 88 bytes of vectors per row.  x is not
// read inside the loop, and K3 applies the same expression to the same p
// before overwriting it, so every x value is the reference's.  F1/F2
// (deterministic fixed-order reductions) update alpha, rho, beta, the
// residual history and a device `done` flag; every kernel returns
// immediately once `done` is set, so a captured chunk of iterations that
// overshoots convergence costs only empty launches and the reported
// iteration count is exact.  The iteration that sets `done` skips its K3;
// its x update is applied by one kernel after the loop (`xpend`).  Chunks of
// iterations (two lengths, 4 and 16) are captured into HIP graphs and
// replayed; the host reads the flag once per chunk with a chunk in flight,
// choosing the next length from the iterations the residual contraction so
// far predicts are left.
#include "fsx_spmv_kernels.hpp"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace fsx {

struct cg_state {
	double rho0, rho1, alpha, beta, res, tol, pw;
	int iter, done, status, iters, maxiter, ncurv;
	int xpend; // the last iteration's x += alpha p is still to be applied
};

// ---- K1: w = A p and the partials of w.p; F1: alpha ----------------------
//
// The row's own p_i comes from the SpMV's gather of the diagonal entry when
// the row has one (no second read of p).  Measured alternative, not used:
// folding K3's p = beta p + z into this gather (two gathered streams) was
// slower than the separate pass on the 256^3 problem.

template<class ACC, class T>
struct epi_cg_spmv {
	static constexpr int NV = 1;
	static constexpr bool reduces = true;
	T * w;
	double * partials;
	const cg_state * st;
	__device__ bool skip() const { return st->done != 0; }
	template<class XI>
	__device__ void row(int64_t i, ACC s, double (&d)[NV], XI && pi) const {
		// Loads before the store: the store would otherwise sit in front of
		// the load in the wave's memory counter and be waited for.
		const double pv = (double)pi();
		const T wi = (T)s;
		w[i] = wi;
		d[0] += (double)wi * pv; // w.dot(p)
	}
	__device__ void block_end(const double (&d)[NV], unsigned blk,
	                          unsigned) const {
		store_block_partials<NV>(d, partials, blk);
	}
};

// Row-sum type of K1: FP32 when matrix and vectors are FP32 (the product
// compressed_ops::spmv computes for a float instantiation), else FP64.
template<class T, class V>
using cg_acc_t = std::conditional_t<std::is_same_v<T, float> && std::is_same_v<V, float>,
                                    float, double>;

// alpha = rho / (w.p)  (cg.hh:98-106; negative curvature is only logged)
struct fin_alpha {
	cg_state * st;
	__device__ bool skip() const { return st->done != 0; }
	__device__ void operator()(const double (&t)[1]) const {
		const double pw = t[0];
		st->pw = pw;
		if (pw <= 0.0) st->ncurv++;
		st->alpha = st->rho1 / pw;
	}
};

// ---- K2: r update, partials of |r|^2 and r.z; F2: test and beta ---------

constexpr int kUpdThreads = 256;
constexpr int kUpdPerThread = 4;

template<class T, bool JAC>
__global__ __launch_bounds__(kUpdThreads) void cg_update_kernel(
	size_t n, T * __restrict__ r, const T * __restrict__ w,
	const T * __restrict__ dinv, T * __restrict__ z, double * partials,
	const cg_state * st) {
	if (st->done) return;
	const double alpha = st->alpha;
	const unsigned blk = blockIdx.x;
	const size_t base = (size_t)blk * kUpdThreads * kUpdPerThread + threadIdx.x;
	double rr = 0.0, rz = 0.0;
#pragma unroll
	for (int j = 0; j < kUpdPerThread; ++j) {
		const size_t i = base + (size_t)j * kUpdThreads;
		if (i < n) {
			// r.axpy(-alpha, w, r)  (cg.hh:109)
			if constexpr (std::is_same_v<T, double>) {
				const double ri = (-alpha) * w[i] + r[i];
				r[i] = ri;
				rr += ri * ri;
				if constexpr (JAC) {
					const double zi = 0.0 + dinv[i] * ri;
					z[i] = zi;
					rz += ri * zi;
				}
			}
			else {
				// FP32 storage, FP64 arithmetic and reductions
				const T ri = (T)((-alpha) * (double)w[i] + (double)r[i]);
				r[i] = ri;
				rr += (double)ri * (double)ri;
				if constexpr (JAC) {
					const T zi = (T)((double)dinv[i] * (double)ri);
					z[i] = zi;
					rz += (double)ri * (double)zi;
				}
			}
		}
	}
	double d[2] = {rr, rz};
	store_block_partials<2>(d, partials, blk);
}

// |r| = sqrt(r.r); diagnostic-free tolerance test; rho and beta
// (cg.hh:111-129)
// XCUR: x was already updated in this iteration (the diagnostic loop), so
// no x update is left pending for after the loop.
template<bool JAC, bool XCUR = false>
struct fin_beta {
	cg_state * st;
	double * hist;
	__device__ bool skip() const { return st->done != 0; }
	__device__ void operator()(const double (&t)[2]) const {
		const double res = sqrt(t[0]);
		st->res = res;
		const int it = st->iter;
		if (hist) hist[it] = res;
		st->iter = it + 1;
		if (res < st->tol) {
			st->iters = it + 1;
			st->status = FSX_CONVERGED_RTOL;
			st->done = 1;
			st->xpend = XCUR ? 0 : 1;
			return;
		}
		if (it + 1 >= st->maxiter) { // loop exhausted
			st->done = 1;
			st->xpend = XCUR ? 0 : 1;
			return;
		}
		// P.apply(r, z) happened in K2; rho and beta (cg.hh:124-129)
		st->rho0 = st->rho1;
		st->rho1 = JAC ? t[1] : t[0];
		st->beta = st->rho1 / st->rho0;
	}
};

// ---- K3: x += alpha p; p = beta p + z  (cg.hh:108, 130) -------------------

template<class T>
__global__ __launch_bounds__(256) void cg_xp_update_kernel(
	size_t n, T * __restrict__ x, T * __restrict__ p, const T * __restrict__ z,
	const cg_state * st) {
	if (st->done) return;
	const double alpha = st->alpha, beta = st->beta;
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) {
		if constexpr (std::is_same_v<T, double>) {
			const double pi = p[i];
			x[i] = alpha * pi + x[i];
			p[i] = beta * pi + z[i];
		}
		else {
			const double pi = (double)p[i];
			x[i] = (T)(alpha * pi + (double)x[i]);
			p[i] = (T)(beta * pi + (double)z[i]);
		}
	}
}

// Diagnostic loop: x += alpha p (cg.hh:108) ahead of F2, since the user
// diagnostic reads x after the update; p = beta p + z (cg.hh:130) after the
// host's check.  The expressions are K3's.
template<class T>
__global__ __launch_bounds__(256) void cg_x_update_kernel(size_t n, T * __restrict__ x,
                                                          const T * __restrict__ p,
                                                          const cg_state * st) {
	if (st->done) return;
	const double alpha = st->alpha;
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) {
		if constexpr (std::is_same_v<T, double>) x[i] = alpha * p[i] + x[i];
		else x[i] = (T)(alpha * (double)p[i] + (double)x[i]);
	}
}

template<class T>
__global__ __launch_bounds__(256) void cg_p_update_kernel(size_t n, T * __restrict__ p,
                                                          const T * __restrict__ z,
                                                          const cg_state * st) {
	if (st->done) return;
	const double beta = st->beta;
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) {
		if constexpr (std::is_same_v<T, double>) p[i] = beta * p[i] + z[i];
		else p[i] = (T)(beta * (double)p[i] + (double)z[i]);
	}
}

// The final iteration's x += alpha p (its K3 was skipped by `done`).
template<class T>
__global__ __launch_bounds__(256) void cg_xfinal_kernel(size_t n, T * __restrict__ x,
                                                        const T * __restrict__ p,
                                                        const cg_state * st) {
	if (!st->xpend) return;
	const double alpha = st->alpha;
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) {
		if constexpr (std::is_same_v<T, double>) x[i] = alpha * p[i] + x[i];
		else x[i] = (T)(alpha * (double)p[i] + (double)x[i]);
	}
}

// ---- single-reduction iteration (unpreconditioned inner solves) -----------
//
// The refinement's inner CG on the symmetrically scaled FP32 matrix has no
// preconditioner (z = r) and no reference counterpart, so it may use the
// single-reduction form of the same recurrences: with r_k . A p_k = p_k . A p_k
// (A-conjugacy), |r_{k+1}|^2 = alpha_k^2 |A p_k|^2 - |r_k|^2.  K1 also sums
// w.w, so after its one finalize alpha_k, that estimate and beta_k are known
// and the r, x and p updates fuse into one pass (28 instead of 32 bytes of
// vectors per row in FP32, one finalize launch fewer).  The pass also sums
// the recursive |r_{k+1}|^2 itself, which the next finalize uses for
// alpha_{k+1} and for the tolerance test of r_{k+1}, so the stopping test is
// the standard one on the recursive residual, taken one kernel later (one
// product after convergence is discarded).

// K1s: w = A p and the partials of w.p and w.w
template<class ACC, class T>
struct epi_cg_spmv2 {
	static constexpr int NV = 2;
	static constexpr bool reduces = true;
	T * w;
	double * partials;
	const cg_state * st;
	__device__ bool skip() const { return st->done != 0; }
	template<class XI>
	__device__ void row(int64_t i, ACC s, double (&d)[NV], XI && pi) const {
		const double pv = (double)pi();
		const T wi = (T)s;
		w[i] = wi;
		d[0] += (double)wi * pv;          // w.dot(p)
		d[1] += (double)wi * (double)wi;  // w.dot(w)
	}
	__device__ void block_end(const double (&d)[NV], unsigned blk, unsigned) const {
		store_block_partials<NV>(d, partials, blk);
	}
};

// F1s: rho_k = |r_k|^2 (the previous fused pass's partials; the setup's rho
// at k = 0) with r_k's tolerance test and maxiter, then alpha_k and beta_k.
__global__ __launch_bounds__(kFinThreads) void sr_finalize_kernel(
	const double * part1, unsigned nblk1, const double * part2, unsigned nblk2,
	cg_state * st, double * hist) {
	if (st->done) return;
	const int it = st->iter; // completed updates
	double t[2];
	sum_partials<2>(part1, nblk1, t);
	double u[1] = {0.0};
	if (it > 0) sum_partials<1>(part2, nblk2, u);
	if (threadIdx.x != 0) return;
	const double rho = it > 0 ? u[0] : st->rho1;
	if (it > 0) {
		const double res = sqrt(rho);
		st->res = res;
		if (hist) hist[it - 1] = res;
		if (res < st->tol) {
			st->iters = it;
			st->status = FSX_CONVERGED_RTOL;
			st->done = 1;
			return;
		}
		if (it >= st->maxiter) { // loop exhausted
			st->done = 1;
			return;
		}
	}
	const double pw = t[0];
	st->pw = pw;
	if (pw <= 0.0) st->ncurv++;
	const double alpha = rho / pw;
	const double rho_next = alpha * alpha * t[1] - rho;
	st->alpha = alpha;
	st->rho0 = rho;
	st->rho1 = rho_next;
	st->beta = rho_next / rho;
	st->iter = it + 1;
}

// K23: r -= alpha w; x += alpha p; p = beta p + r; partials of |r|^2 (the
// expressions of cg_update_kernel and cg_xp_update_kernel)
template<class T>
__global__ __launch_bounds__(kUpdThreads) void cg_sr_update_kernel(
	size_t n, T * __restrict__ x, T * __restrict__ r, T * __restrict__ p,
	const T * __restrict__ w, double * partials, const cg_state * st) {
	if (st->done) return;
	const double alpha = st->alpha, beta = st->beta;
	const unsigned blk = blockIdx.x;
	const size_t base = (size_t)blk * kUpdThreads * kUpdPerThread + threadIdx.x;
	double rr = 0.0;
#pragma unroll
	for (int j = 0; j < kUpdPerThread; ++j) {
		const size_t i = base + (size_t)j * kUpdThreads;
		if (i < n) {
			if constexpr (std::is_same_v<T, double>) {
				const double pi = p[i];
				const double ri = (-alpha) * w[i] + r[i];
				r[i] = ri;
				rr += ri * ri;
				x[i] = alpha * pi + x[i];
				p[i] = beta * pi + ri;
			}
			else {
				const double pi = (double)p[i];
				const T ri = (T)((-alpha) * (double)w[i] + (double)r[i]);
				r[i] = ri;
				rr += (double)ri * (double)ri;
				x[i] = (T)(alpha * pi + (double)x[i]);
				p[i] = (T)(beta * pi + (double)ri);
			}
		}
	}
	double d[1] = {rr};
	store_block_partials<1>(d, partials, blk);
}

// ---- residual + norm (FP64 outer refinement) --------------------------------

template<class ACC, class T>
struct epi_res_norm {
	static constexpr int NV = 1;
	static constexpr bool reduces = true;
	T * r;
	const T * b;
	double * partials;
	__device__ bool skip() const { return false; }
	template<class XI>
	__device__ void row(int64_t i, ACC s, double (&d)[NV], XI &&) const {
		const T ri = b[i] - (T)s;
		r[i] = ri;
		d[0] += (double)ri * (double)ri;
	}
	__device__ void block_end(const double (&d)[NV], unsigned blk,
	                          unsigned) const {
		store_block_partials<NV>(d, partials, blk);
	}
};

// x += s * d (FP64 x, FP32 d) with s read from device memory (|r|)
__global__ void ir_update_kernel(size_t n, double * __restrict__ x,
                                 const float * __restrict__ d,
                                 const double * rnorm2) {
	const double s = sqrt(*rnorm2);
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) x[i] = x[i] + s * (double)d[i];
}

// symmetric Jacobi scaling: r32 = (float)(s (r / |r|)), x += |r| (s d)
__global__ void ir_scale_sym_kernel(size_t n, const double * __restrict__ r,
                                    const double * __restrict__ sc, float * __restrict__ r32,
                                    const double * rnorm2) {
	const double inv = 1.0 / sqrt(*rnorm2);
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) r32[i] = (float)(sc[i] * (r[i] * inv));
}

__global__ void ir_update_sym_kernel(size_t n, double * __restrict__ x,
                                     const float * __restrict__ d,
                                     const double * __restrict__ sc, const double * rnorm2) {
	const double s = sqrt(*rnorm2);
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) x[i] = x[i] + s * (sc[i] * (double)d[i]);
}

// r32 = (float)(r / |r|)
__global__ void ir_scale_kernel(size_t n, const double * __restrict__ r,
                                float * __restrict__ r32,
                                const double * rnorm2) {
	const double inv = 1.0 / sqrt(*rnorm2);
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) r32[i] = (float)(r[i] * inv);
}

// ---- FP32 CG with residual replacement (fsx_ir_cg_rr_solve) ---------------
//
// One FP32 CG on the scaled system S A S y = S r0 / c (c = |r0|) whose
// recursive residual rh is replaced, at the end of each segment, by the true
// one: x += c S y in FP64, r = b - A x in FP64, rh <- (float)(S r / c), and
// p += rh_new - rh_old, so that p = rh_new + beta p_prev as if the new
// residual had been the recursive one (residual replacement, van der Vorst
// and Ye, SIAM J. Sci. Comput. 22 (2000) 835-852).  Unlike a restarted
// refinement the search direction survives, so the Krylov space built so
// far is not thrown away.

// FIRST: rh = (float)(S r / c), p = rh, y = 0; otherwise the replacement.
// Partials of |rh|^2 per block, in the fused update's block layout.
template<bool FIRST>
__global__ __launch_bounds__(kUpdThreads) void rr_replace_kernel(
	size_t n, const double * __restrict__ r, const double * __restrict__ sc, double inv_c,
	float * __restrict__ rh, float * __restrict__ p, float * __restrict__ y,
	double * partials) {
	const unsigned blk = blockIdx.x;
	const size_t base = (size_t)blk * kUpdThreads * kUpdPerThread + threadIdx.x;
	double rr = 0.0;
#pragma unroll
	for (int j = 0; j < kUpdPerThread; ++j) {
		const size_t i = base + (size_t)j * kUpdThreads;
		if (i < n) {
			const float v = (float)(sc[i] * (r[i] * inv_c));
			if constexpr (FIRST) p[i] = v;
			else p[i] = (float)((double)p[i] + ((double)v - (double)rh[i]));
			rh[i] = v;
			y[i] = 0.0f;
			rr += (double)v * (double)v;
		}
	}
	double d[1] = {rr};
	store_block_partials<1>(d, partials, blk);
}

// x += c (S y) in FP64
__global__ void rr_update_kernel(size_t n, double * __restrict__ x, const float * __restrict__ y,
                                 const double * __restrict__ sc, double c) {
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) x[i] = x[i] + c * (sc[i] * (double)y[i]);
}

// What a cached graph's launches touch, for the check before each replay:
// every device range it captured, with its size at capture, and the residual
// history it writes (one entry per iteration up to the solve's maxiter).
namespace graph_ledger {
struct entry {
	std::vector<ledger::span> spans;
	const double * hist = nullptr;
};

// Throws (FSX_ERR_STATE) when a captured range is no longer inside a live
// buffer of at least its captured size.
static void check(const entry & e, int maxiter) {
	auto fail = [](const void * p, size_t b) {
		char msg[160];
		snprintf(msg, sizeof msg,
		         "captured graph refers to %zu bytes at %p that are not inside a live "
		         "buffer; the graph is not launched",
		         b, p);
		throw error(FSX_ERR_STATE, msg);
	};
	for (const auto & s : e.spans)
		if (s.p && s.bytes && !ledger::live(s.p, s.bytes)) fail(s.p, s.bytes);
	if (e.hist && maxiter > 0 && !ledger::live(e.hist, (size_t)maxiter * sizeof(double)))
		fail(e.hist, (size_t)maxiter * sizeof(double));
}

// every array of A's layouts, at its present size
static void add_layout(entry & e, const fsx_csr * A) {
	auto add = [&](const std::unique_ptr<dbuf> & b) {
		if (b && b->p) e.spans.push_back({b->p, b->bytes});
	};
	const csr_index & X = *A->idx;
	add(X.rowptr);
	add(X.colind);
	add(A->vals);
	add(X.sl_ptr);
	add(X.sl_col);
	add(A->sl_vals);
	add(X.dg_ptr);
	add(X.dg_off);
	add(X.dg_mask);
	add(A->dg_vals);
	add(X.st_mask);
	add(A->st_vals);
}
} // namespace graph_ledger

} // namespace fsx

// Per-context solver cache: pooled work vectors (so their addresses are stable
// from one solve to the next), the device state and history buffers, and the
// captured iteration chunks keyed by everything a captured launch depends on.
struct cg_graph_cache {
	struct entry {
		hipGraph_t g = nullptr;
		hipGraphExec_t ge = nullptr;
		fsx::graph_ledger::entry led;
	};
	// key: the iteration function, every device address a captured launch uses and the
	// sizes and plan its launch shapes follow from
	std::map<std::vector<const void *>, entry> mv;
	std::map<std::tuple<size_t, int, int>, fsx_vec *> pool;
	fsx::dbuf * statebuf = nullptr;
	fsx::dbuf * histbuf = nullptr;
	size_t hist_cap = 0;
	void clear_graphs() {
		for (auto & kv : mv) {
			if (kv.second.ge) (void)hipGraphExecDestroy(kv.second.ge);
			if (kv.second.g) (void)hipGraphDestroy(kv.second.g);
		}
		mv.clear();
	}
	~cg_graph_cache() {
		clear_graphs();
		for (auto & kv : pool) fsx_vec_release(kv.second);
		delete statebuf;
		delete histbuf;
	}
};

void cg_cache_deleter::operator()(cg_graph_cache * c) const { delete c; }

namespace fsx {
// Appends every device array a product on A may read (all layouts), and the
// sizes and plan its launch shapes follow from.  A graph keyed with them is
// replayed only when each address it captured belongs to the matrix of the
// current call.
static void append_layout(std::vector<const void *> & key, const fsx_csr * A) {
	auto p = [](const std::unique_ptr<dbuf> & b) -> const void * { return b ? b->p : nullptr; };
	const csr_index & X = *A->idx;
	key.insert(key.end(), {p(X.rowptr), p(X.colind), p(A->vals), p(X.sl_ptr), p(X.sl_col),
	                       p(A->sl_vals), p(X.dg_ptr), p(X.dg_off), p(X.dg_mask),
	                       p(A->dg_vals), p(X.st_mask), p(A->st_vals),
	                       (const void *)(intptr_t)A->nrows,
	                       (const void *)(intptr_t)A->nnz,
	                       (const void *)(intptr_t)A->plan_kind,
	                       (const void *)(intptr_t)A->plan_lanes});
}
} // namespace fsx

void fsx::invalidate_graphs(fsx_ctx * ctx) {
	if (ctx->cg_cache) ctx->cg_cache->clear_graphs();
}

namespace fsx {

struct cg_bufs {
	const fsx_csr * A;
	void * x;
	void * r;
	void * z; // == r for identity
	void * p;
	void * w;
	const void * dinv;
	double * part;  // per-block partials of the reducing kernels
	double * part2; // single-reduction iteration: the fused update's partials
	size_t n;
};

// Device scalar slots of a CG solve's setup reductions (|b|^2, |x|^2, |r|^2,
// rho, final |x|^2) and where the host reads them (pinned, past the state
// copies of the iteration loop).
constexpr int kCgSetupSlot = 16;
constexpr int kCgHostSlot = 40;
static_assert(kCgSetupSlot + 5 <= kScalarSlots && kCgHostSlot + 5 <= kScalarSlots, "");
static_assert(sizeof(cg_state) * 3 <= sizeof(double) * kCgHostSlot,
              "setup values must not overlap the loop's pinned state copies");

// Where a CG solve keeps its work.  The allocating entry points point these
// at context-owned buffers (pooled vectors, the context's partials, the
// cache's state buffer); fsx_cg_solve_noalloc points every device field
// into its static scratch array.
struct cg_work {
	void *r, *z, *p, *w; // z: null without a preconditioner
	double * part;       // partials: at least cg_partials_need(A, n) doubles
	double * slots;      // 5 device scalars (setup reductions, final |x|^2)
	cg_state * st;       // device iteration state
	double * hist;       // device residual history (maxiter entries) or null
	// false: kernels launched directly with one chunk of iterations in
	// flight -- no graphs, events or allocations
	bool graphs;
	// the single-reduction iteration (unpreconditioned solves of the
	// refinement only): K1 with w.p and w.w, one finalize, one fused update
	bool single_reduction = false;
};

// Partials a CG solve on A over n rows needs: the SpMV+dot (1 per block),
// the update (2 per block) and the setup reductions.
static size_t cg_partials_need(const fsx_csr * A, size_t n, bool sr = false) {
	const size_t per = (size_t)kUpdThreads * kUpdPerThread;
	const size_t upd = (n + per - 1) / per;
	size_t need = std::max<size_t>(std::max<size_t>(spmv_partials_size(A, 1), 2 * upd),
	                               (size_t)red_grid(n));
	// single reduction: the product's two sums per block, then the fused
	// update's one per block in a region of its own
	if (sr) need = std::max(need, spmv_partials_size(A, 2) + upd);
	return need;
}

// The ranges a captured CG chunk touches (vectors, partials, state, every
// layout array of A) and its history buffer.
template<class T>
static graph_ledger::entry cg_ledger_entry(const cg_bufs & B, const cg_state * st,
                                           const double * hist) {
	graph_ledger::entry e;
	const size_t vb = B.n * sizeof(T);
	e.spans = {{B.x, vb},
	           {B.r, vb},
	           {B.z, vb},
	           {B.p, vb},
	           {B.w, vb},
	           {B.dinv, B.dinv ? vb : 0},
	           {B.part, cg_partials_need(B.A, B.n, B.part2 != nullptr) * sizeof(double)},
	           {st, sizeof(cg_state)}};
	graph_ledger::add_layout(e, B.A);
	e.hist = hist;
	return e;
}

// One CG iteration (K1, F1, K2, F2, K3).  DIAG: the diagnostic loop's
// iteration -- K1, F1, K2, the x update, F2; the p update is launched by the
// loop after the host has called the diagnostic.
template<class T, class V, class I, bool DIAG = false>
static void enqueue_iteration(fsx_ctx * ctx, const cg_bufs & B, cg_state * st,
                              double * hist, int) {
	hipStream_t s = ctx->stream;
	const fsx_csr * A = B.A;
	double * part = B.part;
	T * p = (T *)B.p;
	// K1 + F1
	using ACC = cg_acc_t<T, V>;
	epi_cg_spmv<ACC, T> e1{(T *)B.w, part, st};
	spmv_launch_typed<decltype(e1), ACC, V, I, T>(s, A, (const T *)p, e1);
	spmv_finalize<1>(s, A, part, fin_alpha{st});
	// K2 + F2
	const size_t per = (size_t)kUpdThreads * kUpdPerThread;
	const unsigned g2 = (unsigned)((B.n + per - 1) / per);
	const unsigned g3 = (unsigned)((B.n + 255) / 256);
	auto xupd = [&] {
		if constexpr (DIAG)
			hipLaunchKernelGGL((cg_x_update_kernel<T>), dim3(g3), dim3(256), 0, s, B.n,
			                   (T *)B.x, (const T *)p, (const cg_state *)st);
	};
	if (B.dinv) {
		hipLaunchKernelGGL((cg_update_kernel<T, true>), dim3(g2), dim3(kUpdThreads),
		                   0, s, B.n, (T *)B.r, (const T *)B.w, (const T *)B.dinv,
		                   (T *)B.z, part, (const cg_state *)st);
		xupd();
		hipLaunchKernelGGL((finalize_kernel<2, fin_beta<true, DIAG>>), dim3(1),
		                   dim3(kFinThreads), 0, s, (const double *)part, g2,
		                   fin_beta<true, DIAG>{st, hist});
	}
	else {
		hipLaunchKernelGGL((cg_update_kernel<T, false>), dim3(g2), dim3(kUpdThreads),
		                   0, s, B.n, (T *)B.r, (const T *)B.w, (const T *)nullptr,
		                   (T *)nullptr, part, (const cg_state *)st);
		xupd();
		hipLaunchKernelGGL((finalize_kernel<2, fin_beta<false, DIAG>>), dim3(1),
		                   dim3(kFinThreads), 0, s, (const double *)part, g2,
		                   fin_beta<false, DIAG>{st, hist});
	}
	// K3
	if constexpr (!DIAG)
		hipLaunchKernelGGL((cg_xp_update_kernel<T>), dim3(g3), dim3(256), 0, s, B.n,
		                   (T *)B.x, p, (const T *)B.z, (const cg_state *)st);
	FSX_HIP(hipGetLastError());
}

// One single-reduction iteration (K1s, F1s, K23); unpreconditioned only.
template<class T, class V, class I>
static void enqueue_iteration_sr(fsx_ctx * ctx, const cg_bufs & B, cg_state * st,
                                 double * hist, int) {
	hipStream_t s = ctx->stream;
	const fsx_csr * A = B.A;
	using ACC = cg_acc_t<T, V>;
	epi_cg_spmv2<ACC, T> e1{(T *)B.w, B.part, st};
	spmv_launch_typed<decltype(e1), ACC, V, I, T>(s, A, (const T *)B.p, e1);
	const size_t per = (size_t)kUpdThreads * kUpdPerThread;
	const unsigned g2 = (unsigned)((B.n + per - 1) / per);
	hipLaunchKernelGGL(sr_finalize_kernel, dim3(1), dim3(kFinThreads), 0, s,
	                   (const double *)B.part, spmv_blocks(A), (const double *)B.part2, g2, st,
	                   hist);
	hipLaunchKernelGGL((cg_sr_update_kernel<T>), dim3(g2), dim3(kUpdThreads), 0, s, B.n,
	                   (T *)B.x, (T *)B.r, (T *)B.p, (const T *)B.w, B.part2,
	                   (const cg_state *)st);
	FSX_HIP(hipGetLastError());
}

// chunk length used while a solve's length is unknown or nearly reached
constexpr int kShortChunk = 4;

static int chunk_size() {
	static int c = [] {
		const char * e = getenv("FSX_CG_CHUNK");
		int v = e ? atoi(e) : 16;
		return v < 1 ? 1 : (v > 256 ? 256 : v);
	}();
	return c;
}

static bool use_graphs() {
	static bool g = [] {
		const char * e = getenv("FSX_CG_GRAPH");
		return !(e && e[0] == '0');
	}();
	return g;
}

static cg_graph_cache & cache_of(fsx_ctx * ctx) {
	if (!ctx->cg_cache) ctx->cg_cache.reset(new cg_graph_cache());
	auto & c = *ctx->cg_cache;
	if (!c.statebuf) c.statebuf = new dbuf(sizeof(cg_state) * 4);
	return c;
}

// Device residual-history buffer with room for `maxiter` entries.  Growing
// it invalidates captured graphs (they hold its address).
static double * history_buffer(fsx_ctx * ctx, int maxiter) {
	auto & c = cache_of(ctx);
	size_t need = (size_t)std::max(1, maxiter);
	if (need > c.hist_cap) {
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		c.clear_graphs();
		delete c.histbuf;
		c.histbuf = nullptr;
		size_t cap = std::max<size_t>(need, 4096);
		c.histbuf = new dbuf(cap * sizeof(double));
		c.hist_cap = cap;
	}
	return (double *)c.histbuf->p;
}

// Pooled work vector `slot` of length n and dtype dt (owned by the context).
static fsx_vec * pooled(fsx_ctx * ctx, size_t n, fsx_dtype dt, int slot) {
	auto & c = cache_of(ctx);
	auto key = std::make_tuple(n, (int)dt, slot);
	auto it = c.pool.find(key);
	if (it != c.pool.end()) return it->second;
	fsx_vec * v = nullptr;
	int rc = fsx_vec_create(ctx, n, dt, &v);
	if (rc) throw error(rc, get_error());
	c.pool[key] = v;
	return v;
}

// Work of an allocating CG solve: pooled context vectors in slots slot0 ..
// slot0 + 3 (r, z, p, w), the context's partials grown to what the solve
// needs, the cache's state buffer and a history buffer for maxiter entries.
static cg_work pooled_work(fsx_ctx * ctx, const fsx_csr * A, size_t n, fsx_dtype dt,
                           bool jac, int slot0, int maxiter, bool sr = false) {
	cg_work wk{};
	wk.r = pooled(ctx, n, dt, slot0 + 0)->ptr();
	wk.z = jac ? pooled(ctx, n, dt, slot0 + 1)->ptr() : nullptr;
	wk.p = pooled(ctx, n, dt, slot0 + 2)->ptr();
	wk.w = pooled(ctx, n, dt, slot0 + 3)->ptr();
	wk.single_reduction = sr && !jac;
	wk.part = ensure_partials(ctx, cg_partials_need(A, n, wk.single_reduction));
	wk.slots = ctx->ws.scalars + kCgSetupSlot;
	wk.st = (cg_state *)cache_of(ctx).statebuf->p;
	wk.hist = maxiter > 0 ? history_buffer(ctx, maxiter) : nullptr;
	wk.graphs = true;
	return wk;
}

// Iteration loop without graphs, events or allocations (fsx_cg_solve_noalloc):
// chunks of iterations launched directly, the host reading the state after
// each.  Same kernels in the same order as the captured loop, so the same
// iterates.
template<class T, class V, class I>
static void run_cg_loop_direct(fsx_ctx * ctx, const cg_bufs & B, cg_state * st,
                               cg_state * pinned, double * hist) {
	const int K = chunk_size();
	const long maxit = pinned->maxiter;
	cg_state * flag = pinned + 1;
	long launched_it = 0;
	while (launched_it < maxit) {
		const int len = (int)std::min<long>(K, maxit - launched_it);
		for (int k = 0; k < len; ++k) enqueue_iteration<T, V, I>(ctx, B, st, hist, k);
		ctx->launches += 5 * len;
		launched_it += len;
		FSX_HIP(hipMemcpyAsync(flag, st, sizeof(cg_state), hipMemcpyDeviceToHost,
		                       ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		ctx->host_syncs++;
		if (flag->done) break;
	}
}

// Captured, pipelined loop: chunks of iterations replayed from cached HIP
// graphs (or launched directly with FSX_CG_GRAPH=0), up to two in flight,
// the host checking the older one's state.
template<class T, class V, class I, bool SR = false>
static void run_cg_loop_pipelined(fsx_ctx * ctx, const cg_bufs & B, cg_state * st,
                                  cg_state * pinned, double * hist,
                                  const cg_state & st0) {
	// the iteration: the reference's sequence, or the single-reduction form
	constexpr auto iterate = SR ? &enqueue_iteration_sr<T, V, I> : &enqueue_iteration<T, V, I>;
	constexpr int kLaunches = SR ? 3 : 5;
	const int K = chunk_size();
	// Two chunk lengths: short chunks while the solve's length is unknown or
	// nearly reached, long ones in between.  Iterations past convergence are
	// launches that return at once, so the schedule changes only how many of
	// those run, never the iterates.
	const int Kbig = K;
	const int Ksmall = std::min(Kbig, kShortChunk);
	hipGraphExec_t gexec[2] = {nullptr, nullptr}; // [0] Ksmall, [1] Kbig
	const graph_ledger::entry * gled[2] = {nullptr, nullptr};
	if (use_graphs()) {
		auto & cache = cache_of(ctx);
		auto key_for = [&](int k) {
			std::vector<const void *> key{(const void *)iterate,
			                              B.x, B.r, B.z, B.p, B.w, B.dinv,
			                              (const void *)B.part, (const void *)st,
			                              (const void *)hist, (const void *)(intptr_t)B.n,
			                              (const void *)(intptr_t)k};
			append_layout(key, B.A);
			return key;
		};
		const int lens[2] = {Ksmall, Kbig};
		const int nlen = Ksmall == Kbig ? 1 : 2;
		std::vector<const void *> keys[2];
		int missing = 0;
		for (int j = 0; j < nlen; ++j) {
			keys[j] = key_for(lens[j]);
			missing += cache.mv.count(keys[j]) ? 0 : 1;
		}
		// make room before capturing, so no executable obtained here is freed
		// (the previous solve ended with a stream synchronisation)
		if (missing && cache.mv.size() + missing > 32) cache.clear_graphs();
		for (int j = 0; j < nlen; ++j) {
			auto it = cache.mv.find(keys[j]);
			if (it == cache.mv.end()) {
				cg_graph_cache::entry en;
				FSX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
				try {
					for (int k = 0; k < lens[j]; ++k)
						iterate(ctx, B, st, hist, k);
				}
				catch (...) {
					hipGraph_t junk = nullptr;
					(void)hipStreamEndCapture(ctx->stream, &junk);
					if (junk) (void)hipGraphDestroy(junk);
					throw;
				}
				FSX_HIP(hipStreamEndCapture(ctx->stream, &en.g));
				FSX_HIP(hipGraphInstantiate(&en.ge, en.g, nullptr, nullptr, 0));
				en.led = cg_ledger_entry<T>(B, st, hist);
				it = cache.mv.emplace(keys[j], std::move(en)).first;
			}
			gexec[j] = it->second.ge;
			gled[j] = &it->second.led;
		}
		if (nlen == 1) {
			gexec[1] = gexec[0];
			gled[1] = gled[0];
		}
	}
	cg_state * flags = pinned + 1;
	hipEvent_t ev[2];
	FSX_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
	FSX_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
	auto launch_chunk = [&](int slot, int len) {
		const int j = len == Ksmall ? 0 : 1;
		if (gexec[j]) {
			// every address the graph captured must still be a live buffer
			// at least as large as at capture
			graph_ledger::check(*gled[j], st0.maxiter);
			FSX_HIP(hipGraphLaunch(gexec[j], ctx->stream));
		}
		else
			for (int k = 0; k < len; ++k) iterate(ctx, B, st, hist, k);
		ctx->launches += kLaunches * len;
		FSX_HIP(hipMemcpyAsync(&flags[slot], st, sizeof(cg_state),
		                       hipMemcpyDeviceToHost, ctx->stream));
		FSX_HIP(hipEventRecord(ev[slot], ctx->stream));
	};
	const long maxit = st0.maxiter;
	const double res0 = st0.res, tol = st0.tol;
	long launched_it = 0; // iterations launched
	long want = -1;       // iterations predicted beyond those launched (-1: unknown)
	int launched = 0, checked = 0;
	try {
		for (;;) {
			const bool in_flight = launched > checked;
			if (launched - checked < 2 && launched_it < maxit && (!in_flight || want != 0)) {
				const int len =
					(launched < 2 || (want >= 0 && want <= Ksmall + 2)) ? Ksmall : Kbig;
				launch_chunk(launched & 1, len);
				++launched;
				launched_it += len;
				if (want > 0) want = std::max<long>(0, want - len);
			}
			if (checked == launched) break; // nothing in flight and nothing to launch
			FSX_HIP(hipEventSynchronize(ev[checked & 1]));
			ctx->host_syncs++;
			const cg_state & f = flags[checked & 1];
			++checked;
			if (f.done) break;
			// remaining iterations at the mean contraction per iteration so far
			want = -1;
			if (f.iter > 0 && f.res > 0 && res0 > 0 && tol > 0) {
				const double rate = std::pow(f.res / res0, 1.0 / f.iter);
				if (rate > 0 && rate < 1) {
					const double m = std::ceil(std::log(tol / f.res) / std::log(rate));
					if (m < 1e9) want = std::max<long>(0, (long)m - (launched_it - f.iter));
				}
			}
		}
	}
	catch (...) {
		(void)hipStreamSynchronize(ctx->stream);
		(void)hipEventDestroy(ev[0]);
		(void)hipEventDestroy(ev[1]);
		throw;
	}
	(void)hipEventDestroy(ev[0]);
	(void)hipEventDestroy(ev[1]);
}

// Iteration loop with a user diagnostic (cg.hh:112-116): one iteration at a
// time, the host reading |r| after each and calling diag(x, |r|) before the
// tolerance test takes effect -- a true return stops the solve with
// converged_user and iters = it + 1 (also when the tolerance was met in the
// same iteration, as the reference checks the diagnostic first).
template<class T, class V, class I>
static void run_cg_loop_diag(fsx_ctx * ctx, const cg_bufs & B, cg_state * st,
                             cg_state * pinned, double * hist, fsx_vec * xv,
                             fsx_diag_fn fn, void * user) {
	cg_state * flag = pinned + 1;
	const int maxit = pinned->maxiter;
	const unsigned g3 = (unsigned)((B.n + 255) / 256);
	*flag = *pinned;
	for (int it = 0; it < maxit; ++it) {
		enqueue_iteration<T, V, I, true>(ctx, B, st, hist, it);
		ctx->launches += 5;
		FSX_HIP(hipMemcpyAsync(flag, st, sizeof(cg_state), hipMemcpyDeviceToHost,
		                       ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		ctx->host_syncs++;
		const int d = fn(user, xv, flag->res);
		if (d < 0) throw error(FSX_ERR_STATE, "cg: the user diagnostic reported an error");
		if (d > 0) {
			flag->iters = it + 1;
			flag->status = FSX_CONVERGED_USER;
			flag->done = 1;
			break;
		}
		if (flag->done) break;
		if (B.n) {
			hipLaunchKernelGGL((cg_p_update_kernel<T>), dim3(g3), dim3(256), 0, ctx->stream,
			                   B.n, (T *)B.p, (const T *)B.z, (const cg_state *)st);
			FSX_HIP(hipGetLastError());
			ctx->launches++;
		}
	}
	*pinned = *flag;
}

// Runs the iteration loop on the device.  `st_host` holds the initial state
// and receives the final one; *x2_out receives |x|^2 of the final x.
template<class T, class V, class I>
static void run_cg_loop(fsx_ctx * ctx, const cg_bufs & B, const cg_work & wk,
                        cg_state & st_host, fsx_vec * xv, double * x2_out,
                        fsx_diag_fn diag, void * user) {
	cg_state * st = wk.st;
	cg_state * pinned = (cg_state *)ctx->host_scalars; // pinned host space
	static_assert(sizeof(cg_state) * 3 <= sizeof(double) * kScalarSlots, "");
	std::memcpy(pinned, &st_host, sizeof(cg_state));
	FSX_HIP(hipMemcpyAsync(st, pinned, sizeof(cg_state), hipMemcpyHostToDevice,
	                       ctx->stream));
	if (diag) run_cg_loop_diag<T, V, I>(ctx, B, st, pinned, wk.hist, xv, diag, user);
	else if (wk.graphs && B.part2)
		run_cg_loop_pipelined<T, V, I, true>(ctx, B, st, pinned, wk.hist, st_host);
	else if (wk.graphs) run_cg_loop_pipelined<T, V, I>(ctx, B, st, pinned, wk.hist, st_host);
	else run_cg_loop_direct<T, V, I>(ctx, B, st, pinned, wk.hist);
	if (B.n > 0 && !diag) {
		// the iteration that set `done` skipped K3: apply its x update
		const unsigned g3 = (unsigned)((B.n + 255) / 256);
		hipLaunchKernelGGL((cg_xfinal_kernel<T>), dim3(g3), dim3(256), 0, ctx->stream,
		                   B.n, (T *)B.x, (const T *)B.p, (const cg_state *)st);
		FSX_HIP(hipGetLastError());
		ctx->launches++;
	}
	// the final state (the diagnostic loop has it on the host already, with
	// its own stop recorded)
	if (!diag)
		FSX_HIP(hipMemcpyAsync(pinned, st, sizeof(cg_state), hipMemcpyDeviceToHost,
		                       ctx->stream));
	// |x|^2 of the final x, read with the same synchronisation
	double * x2slot = wk.slots + 4;
	if (xv && B.n > 0) {
		reduce_raw(ctx, RED_L2SQ, xv->dt, B.n, xv->ptr(), nullptr, x2slot, wk.part);
		FSX_HIP(hipMemcpyAsync(ctx->host_scalars + kCgHostSlot + 4, x2slot, sizeof(double),
		                       hipMemcpyDeviceToHost, ctx->stream));
	}
	FSX_HIP(hipStreamSynchronize(ctx->stream));
	std::memcpy(&st_host, pinned, sizeof(cg_state));
	if (x2_out) *x2_out = (xv && B.n > 0) ? ctx->host_scalars[kCgHostSlot + 4] : 0.0;
}

// r = b - A x (op::core::residual, operators/core.hh:129-142: apply, then
// r = b - r), the same expression as fsx_csr_residual's epilogue
template<class ACC, class T>
struct epi_cg_residual {
	static constexpr int NV = 1;
	static constexpr bool reduces = false;
	T * r;
	const T * b;
	__device__ bool skip() const { return false; }
	template<class XI>
	__device__ void row(int64_t i, ACC s, double (&)[NV], XI &&) const {
		r[i] = b[i] - (T)s;
	}
	__device__ void block_end(const double (&)[NV], unsigned, unsigned) const {}
};

// z = D^{-1} r as the Dinv CSR row computes it (y_i = 0; y_i += dinv_i r_i),
// the expression of fsx_jacobi_apply
template<class T>
__global__ __launch_bounds__(256) void cg_jacobi_kernel(size_t n, const T * __restrict__ dinv,
                                                        const T * __restrict__ r,
                                                        T * __restrict__ z) {
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) z[i] = T(0) + dinv[i] * r[i];
}

// Shared driver for FP64 (T = double) and FP32 inner (T = float) solves.
// Follows cg.hh:43-139 step by step; see the file comment.  The work vectors,
// partials, scalar slots and state come from `wk` (allocates nothing itself).
// `hist_count` receives the number of residual-history entries written on
// the device.
template<class T>
static void cg_solve_impl(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv,
                          const fsx_vec * b, fsx_vec * x, const fsx_settings & s,
                          fsx_info * info, const cg_work & wk, int * hist_count,
                          fsx_diag_fn diag = nullptr, void * user = nullptr) {
	std::memset(info, 0, sizeof(*info));
	info->status = FSX_UNKNOWN;
	if (hist_count) *hist_count = 0;
	const size_t n = b->n;
	const fsx_dtype dt = b->dt;
	T * r = (T *)wk.r;
	// The setup reductions |b|^2, |x|^2, |r|^2 and rho = z.r are enqueued
	// into device slots and read with one synchronisation (the same kernels
	// as one-by-one reads, so the same values); z = D^{-1} r is formed even
	// when the initial residual already meets the tolerance.
	double * sl = wk.slots;
	double * hs = ctx->host_scalars + kCgHostSlot;
	if (s.use_zero_guess) {
		launch_set(ctx, 0., x);
		if (n)
			FSX_HIP(hipMemcpyAsync(r, b->ptr(), n * sizeof(T), hipMemcpyDeviceToDevice,
			                       ctx->stream));
	}
	else if (n) {
		reduce_raw(ctx, RED_L2SQ, dt, n, x->ptr(), nullptr, sl + 1, wk.part);
		by_matrix_type(A, [&](auto * vtag, auto * itag) {
			using V = std::remove_pointer_t<decltype(vtag)>;
			using I = std::remove_pointer_t<decltype(itag)>;
			using ACC = cg_acc_t<T, V>;
			epi_cg_residual<ACC, T> e{r, (const T *)b->ptr()};
			spmv_launch_typed<decltype(e), ACC, V, I, T>(ctx->stream, A, (const T *)x->ptr(),
			                                             e);
		});
		ctx->launches++;
	}
	T * zz = r; // op::I: z is a copy of r, so read r directly
	if (dinv) {
		zz = (T *)wk.z;
		if (n) {
			hipLaunchKernelGGL((cg_jacobi_kernel<T>), dim3((unsigned)((n + 255) / 256)),
			                   dim3(256), 0, ctx->stream, n, (const T *)dinv->ptr(),
			                   (const T *)r, zz);
			FSX_HIP(hipGetLastError());
			ctx->launches++;
		}
	}
	double b2 = 0, x2 = 0, r2 = 0, rho = 0;
	if (n) {
		reduce_raw(ctx, RED_L2SQ, dt, n, b->ptr(), nullptr, sl + 0, wk.part);
		reduce_raw(ctx, RED_L2SQ, dt, n, r, nullptr, sl + 2, wk.part);
		reduce_raw(ctx, RED_DOT, dt, n, zz, r, sl + 3, wk.part);
		FSX_HIP(hipMemcpyAsync(hs, sl, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		ctx->host_syncs++;
		b2 = hs[0];
		x2 = hs[1];
		r2 = hs[2];
		rho = hs[3];
	}
	double b_norm = std::sqrt(b2);
	if (b_norm == 0.0) b_norm = 1.0;
	const double terminate_tol = s.rtol * b_norm;
	info->rhs_norm = (float)b_norm;
	info->sol_norm_initial = s.use_zero_guess ? 0.f : (float)std::sqrt(x2);
	double current_res = std::sqrt(r2);
	if (current_res < terminate_tol) {
		info->res_norm_initial = (float)current_res;
		info->res_norm_final = (float)current_res;
		info->status = FSX_CONVERGED_RTOL;
		return;
	}
	cg_state st{};
	st.rho1 = rho;
	st.rho0 = st.rho1;
	st.tol = terminate_tol;
	st.res = current_res;
	st.maxiter = s.maxiter;
	if (n)
		FSX_HIP(hipMemcpyAsync(wk.p, zz, n * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
	double xf2 = 0;
	if (s.maxiter > 0) {
		const bool sr = wk.single_reduction && !dinv && !diag;
		cg_bufs B{A, x->ptr(), r, zz, wk.p, wk.w, dinv ? dinv->ptr() : nullptr, wk.part,
		          sr ? wk.part + spmv_partials_size(A, 2) : nullptr, n};
		by_matrix_type(A, [&](auto * vtag, auto * itag) {
			using V = std::remove_pointer_t<decltype(vtag)>;
			using I = std::remove_pointer_t<decltype(itag)>;
			run_cg_loop<T, V, I>(ctx, B, wk, st, x, &xf2, diag, user);
		});
		current_res = st.res;
		info->iters = st.iters;
		if (st.iters) info->status = st.status;
		if (hist_count) *hist_count = st.iter;
	}
	else if (n) {
		reduce_raw(ctx, RED_L2SQ, dt, n, x->ptr(), nullptr, sl + 4, wk.part);
		FSX_HIP(hipMemcpyAsync(hs + 4, sl + 4, sizeof(double), hipMemcpyDeviceToHost,
		                       ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		ctx->host_syncs++;
		xf2 = hs[4];
	}
	info->res_norm_final = (float)current_res;
	info->sol_norm_final = (float)std::sqrt(xf2);
	if (info->iters == 0) info->status = FSX_DIVERGED_ITERS;
}

// ---- standalone fused CG update (cg.hh:108-127 in one pass) -----------------

template<bool JAC>
__global__ __launch_bounds__(kUpdThreads) void cg_update_full_kernel(
	size_t n, double alpha, double * __restrict__ x, double * __restrict__ r,
	const double * __restrict__ p, const double * __restrict__ w,
	const double * __restrict__ dinv, double * __restrict__ z, double * partials) {
	const unsigned blk = blockIdx.x;
	const size_t base = (size_t)blk * kUpdThreads * kUpdPerThread + threadIdx.x;
	double rr = 0.0, rz = 0.0;
#pragma unroll
	for (int j = 0; j < kUpdPerThread; ++j) {
		const size_t i = base + (size_t)j * kUpdThreads;
		if (i < n) {
			x[i] = alpha * p[i] + x[i];
			const double ri = (-alpha) * w[i] + r[i];
			r[i] = ri;
			rr += ri * ri;
			if constexpr (JAC) {
				const double zi = 0.0 + dinv[i] * ri;
				z[i] = zi;
				rz += ri * zi;
			}
		}
	}
	double d[2] = {rr, JAC ? rz : rr};
	store_block_partials<2>(d, partials, blk);
}

// ---- fused BiCGSTAB (op::bicgstab::apply, solvers/bicgstab.hh:39-187) ----
//
// Same operation sequence, tolerances and solve_info as the reference; the
// scalars live on the device.  One iteration:
//   B1  p = beta (p - omega v) + res (res on the first iteration); p^ = P p
//                                                    (bicgstab.hh:114-124)
//   B2  v = A p^, partials of r~.v                  (126-128)   SpMV + dot
//   G2  alpha = rho / (r~.v)                        (128-130)
//   B3  s = res - alpha v, partials of |s|^2; s^ = P s   (132-140)
//   G3  |s| < rtol (absolute, as the reference compares) -> early exit with
//       iters = iter; x += alpha p^ applied by one kernel after the loop
//   B4  t = A s^, partials of t.t and t.s           (142-146)   SpMV + dots
//   G4  omega = t.s / t.t (0 when t.t = 0)
//   B5  x += alpha p^; x += omega s^; res = s - omega t; partials of |res|^2
//       and r~.res (the next iteration's rho)       (148-153, 105)
//   G5  |res| test, omega = 0 breakdown, rho0 = rho, iteration count, the
//       next rho and its breakdown test, beta      (155-183, 105-111, 118)
// A breakdown restart (r~ orthogonal to res, 107-116) ends the chunk; the
// host restarts from the true residual and resumes.

struct bcg_state {
	double rho0, rho1, alpha, omega, beta, res, tol, rtol_abs, rt_norm;
	int iter, done, status, iters, maxiter, first, restart, early;
};

template<bool JAC>
__global__ __launch_bounds__(256) void bcg_p_kernel(size_t n, double * __restrict__ p,
                                                    double * __restrict__ ph,
                                                    const double * __restrict__ v,
                                                    const double * __restrict__ res,
                                                    const double * __restrict__ dinv,
                                                    const bcg_state * st) {
	if (st->done) return;
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i >= n) return;
	double pi;
	if (st->first) pi = res[i];
	else {
		const double t = (-st->omega) * v[i] + p[i]; // p.axpy(-omega, v, p)
		pi = st->beta * t + res[i];                  // p.axpy(beta, p, res)
	}
	p[i] = pi;
	if constexpr (JAC) ph[i] = 0.0 + dinv[i] * pi;
}

template<class ACC, class T>
struct epi_bcg_v {
	static constexpr int NV = 1;
	static constexpr bool reduces = true;
	T * v;
	const T * rt;
	double * partials;
	const bcg_state * st;
	__device__ bool skip() const { return st->done != 0; }
	template<class XI>
	__device__ void row(int64_t i, ACC s, double (&d)[NV], XI &&) const {
		const double rti = (double)rt[i]; // load before the store
		const T vi = (T)s;
		v[i] = vi;
		d[0] += rti * (double)vi; // r_tilde.dot(v)
	}
	__device__ void block_end(const double (&d)[NV], unsigned blk, unsigned) const {
		store_block_partials<NV>(d, partials, blk);
	}
};

struct fin_bcg_alpha {
	bcg_state * st;
	__device__ bool skip() const { return st->done != 0; }
	__device__ void operator()(const double (&t)[1]) const {
		if (t[0] == 0.0) { // the reference asserts alpha != 0
			st->status = FSX_DIVERGED_BREAKDOWN;
			st->done = 1;
			return;
		}
		st->alpha = st->rho1 / t[0];
	}
};

template<bool JAC>
__global__ __launch_bounds__(kUpdThreads) void bcg_s_kernel(
	size_t n, double * __restrict__ s, double * __restrict__ sh,
	const double * __restrict__ v, const double * __restrict__ res,
	const double * __restrict__ dinv, double * partials, const bcg_state * st) {
	if (st->done) return;
	const double alpha = st->alpha;
	const unsigned blk = blockIdx.x;
	const size_t base = (size_t)blk * kUpdThreads * kUpdPerThread + threadIdx.x;
	double ss = 0.0;
#pragma unroll
	for (int j = 0; j < kUpdPerThread; ++j) {
		const size_t i = base + (size_t)j * kUpdThreads;
		if (i < n) {
			const double si = (-alpha) * v[i] + res[i]; // s.axpy(-alpha, v, res)
			s[i] = si;
			ss += si * si;
			if constexpr (JAC) sh[i] = 0.0 + dinv[i] * si;
		}
	}
	double d[1] = {ss};
	store_block_partials<1>(d, partials, blk);
}

struct fin_bcg_s {
	bcg_state * st;
	__device__ bool skip() const { return st->done != 0; }
	__device__ void operator()(const double (&t)[1]) const {
		if (sqrt(t[0]) < st->rtol_abs) { // early convergence, iters = iter
			st->early = 1;
			st->iters = st->iter;
			st->status = FSX_CONVERGED_RTOL;
			st->done = 1;
		}
	}
};

template<class ACC, class T>
struct epi_bcg_t {
	static constexpr int NV = 2;
	static constexpr bool reduces = true;
	T * t;
	const T * s;
	double * partials;
	const bcg_state * st;
	__device__ bool skip() const { return st->done != 0; }
	template<class XI>
	__device__ void row(int64_t i, ACC sum, double (&d)[NV], XI &&) const {
		const double si = (double)s[i]; // load before the store
		const T ti = (T)sum;
		t[i] = ti;
		d[0] += (double)ti * (double)ti; // t.dot(t)
		d[1] += (double)ti * si; // t.dot(s)
	}
	__device__ void block_end(const double (&d)[NV], unsigned blk, unsigned) const {
		store_block_partials<NV>(d, partials, blk);
	}
};

struct fin_bcg_omega {
	bcg_state * st;
	__device__ bool skip() const { return st->done != 0; }
	__device__ void operator()(const double (&t)[2]) const {
		st->omega = t[0] == 0.0 ? 0.0 : t[1] / t[0];
	}
};

__global__ __launch_bounds__(kUpdThreads) void bcg_x_kernel(
	size_t n, double * __restrict__ x, double * __restrict__ res,
	const double * __restrict__ ph, const double * __restrict__ sh,
	const double * __restrict__ t, const double * __restrict__ s,
	const double * __restrict__ rt, double * partials, const bcg_state * st) {
	if (st->done) return;
	const double alpha = st->alpha, omega = st->omega;
	const unsigned blk = blockIdx.x;
	const size_t base = (size_t)blk * kUpdThreads * kUpdPerThread + threadIdx.x;
	double rr = 0.0, rtr = 0.0;
#pragma unroll
	for (int j = 0; j < kUpdPerThread; ++j) {
		const size_t i = base + (size_t)j * kUpdThreads;
		if (i < n) {
			double xi = alpha * ph[i] + x[i]; // x.axpy(alpha, p_hat, x)
			xi = omega * sh[i] + xi;          // x.axpy(omega, s_hat, x)
			x[i] = xi;
			const double ri = (-omega) * t[i] + s[i]; // res.axpy(-omega, t, s)
			res[i] = ri;
			rr += ri * ri;
			rtr += rt[i] * ri; // next rho: r_tilde.dot(res)
		}
	}
	double d[2] = {rr, rtr};
	store_block_partials<2>(d, partials, blk);
}

struct fin_bcg_res {
	bcg_state * st;
	__device__ bool skip() const { return st->done != 0; }
	__device__ void operator()(const double (&t)[2]) const {
		const double res = sqrt(t[0]);
		st->res = res;
		const int it = st->iter;
		if (res < st->tol) {
			st->status = FSX_CONVERGED_RTOL;
			st->iters = it + 1;
			st->done = 1;
			return;
		}
		if (st->omega == 0.0) {
			st->iters = it + 1;
			st->status = FSX_DIVERGED_BREAKDOWN;
			st->done = 1;
			return;
		}
		st->rho0 = st->rho1;
		st->iter = it + 1;
		st->first = 0;
		if (it + 1 >= st->maxiter) { // loop exhausted: iters stays 0
			st->done = 1;
			return;
		}
		// top of the next iteration (bicgstab.hh:105-118)
		const double rho = t[1];
		st->rho1 = rho;
		if (sqrt(fabs(rho)) < DBL_EPSILON * st->rt_norm) {
			st->restart = 1;
			st->done = 1;
			return;
		}
		st->beta = (rho / st->rho0) * (st->alpha / st->omega);
	}
};

__global__ __launch_bounds__(256) void bcg_xearly_kernel(size_t n, double * __restrict__ x,
                                                         const double * __restrict__ ph,
                                                         const bcg_state * st) {
	if (!st->early) return;
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) x[i] = st->alpha * ph[i] + x[i];
}

struct bcg_bufs {
	const fsx_csr * A;
	double *x, *res, *rt, *p, *v, *ph, *s, *sh, *t;
	const double * dinv;
	size_t n;
};

template<class V, class I>
static void bcg_enqueue_iteration(fsx_ctx * ctx, const bcg_bufs & B, bcg_state * st) {
	hipStream_t s = ctx->stream;
	double * part = ctx->ws.partials;
	const unsigned g1 = (unsigned)((B.n + 255) / 256);
	const size_t per = (size_t)kUpdThreads * kUpdPerThread;
	const unsigned gu = (unsigned)((B.n + per - 1) / per);
	const bool jac = B.dinv != nullptr;
	// B1
	if (jac)
		hipLaunchKernelGGL((bcg_p_kernel<true>), dim3(g1), dim3(256), 0, s, B.n, B.p, B.ph,
		                   (const double *)B.v, (const double *)B.res, B.dinv,
		                   (const bcg_state *)st);
	else
		hipLaunchKernelGGL((bcg_p_kernel<false>), dim3(g1), dim3(256), 0, s, B.n, B.p, B.ph,
		                   (const double *)B.v, (const double *)B.res, B.dinv,
		                   (const bcg_state *)st);
	// B2 + G2
	epi_bcg_v<double, double> e2{B.v, B.rt, part, st};
	spmv_launch_typed<decltype(e2), double, V, I, double>(s, B.A, (const double *)B.ph, e2);
	spmv_finalize<1>(s, B.A, part, fin_bcg_alpha{st});
	// B3 + G3
	if (jac)
		hipLaunchKernelGGL((bcg_s_kernel<true>), dim3(gu), dim3(kUpdThreads), 0, s, B.n, B.s,
		                   B.sh, (const double *)B.v, (const double *)B.res, B.dinv, part,
		                   (const bcg_state *)st);
	else
		hipLaunchKernelGGL((bcg_s_kernel<false>), dim3(gu), dim3(kUpdThreads), 0, s, B.n,
		                   B.s, B.sh, (const double *)B.v, (const double *)B.res, B.dinv,
		                   part, (const bcg_state *)st);
	hipLaunchKernelGGL((finalize_kernel<1, fin_bcg_s>), dim3(1), dim3(kFinThreads), 0, s,
	                   (const double *)part, gu, fin_bcg_s{st});
	// B4 + G4
	epi_bcg_t<double, double> e4{B.t, B.s, part, st};
	spmv_launch_typed<decltype(e4), double, V, I, double>(s, B.A, (const double *)B.sh, e4);
	spmv_finalize<2>(s, B.A, part, fin_bcg_omega{st});
	// B5 + G5
	hipLaunchKernelGGL(bcg_x_kernel, dim3(gu), dim3(kUpdThreads), 0, s, B.n, B.x, B.res,
	                   (const double *)B.ph, (const double *)B.sh, (const double *)B.t,
	                   (const double *)B.s, (const double *)B.rt, part,
	                   (const bcg_state *)st);
	hipLaunchKernelGGL((finalize_kernel<2, fin_bcg_res>), dim3(1), dim3(kFinThreads), 0, s,
	                   (const double *)part, gu, fin_bcg_res{st});
	FSX_HIP(hipGetLastError());
}

// Runs device iterations from the state in st_host until the device sets
// `done` (converged, breakdown, early exit, exhausted, or restart needed).
template<class V, class I>
static void bcg_run_chunks(fsx_ctx * ctx, const bcg_bufs & B, bcg_state & st_host) {
	auto & cache = cache_of(ctx);
	bcg_state * st = (bcg_state *)cache.statebuf->p;
	bcg_state * pinned = (bcg_state *)ctx->host_scalars;
	static_assert(sizeof(bcg_state) * 3 <= sizeof(double) * kScalarSlots, "");
	static_assert(sizeof(bcg_state) <= sizeof(cg_state) * 4, "");
	std::memcpy(pinned, &st_host, sizeof(bcg_state));
	FSX_HIP(hipMemcpyAsync(st, pinned, sizeof(bcg_state), hipMemcpyHostToDevice,
	                       ctx->stream));
	const int K = chunk_size();
	{
		const size_t per = (size_t)kUpdThreads * kUpdPerThread;
		ensure_partials(ctx, std::max<size_t>(spmv_partials_size(B.A, 2),
		                                      2 * ((B.n + per - 1) / per)));
	}
	hipGraphExec_t ge = nullptr;
	if (use_graphs()) {
		std::vector<const void *> key{(const void *)&bcg_enqueue_iteration<V, I>,
		                              B.x, B.res, B.rt, B.p, B.v, B.ph, B.s, B.sh, B.t,
		                              B.dinv, (const void *)ctx->ws.partials,
		                              (const void *)st, (const void *)(intptr_t)B.n,
		                              (const void *)(intptr_t)K};
		append_layout(key, B.A);
		auto it = cache.mv.find(key);
		if (it == cache.mv.end()) {
			if (cache.mv.size() >= 32) cache.clear_graphs();
			cg_graph_cache::entry en;
			FSX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
			try {
				for (int k = 0; k < K; ++k) bcg_enqueue_iteration<V, I>(ctx, B, st);
			}
			catch (...) {
				hipGraph_t junk = nullptr;
				(void)hipStreamEndCapture(ctx->stream, &junk);
				if (junk) (void)hipGraphDestroy(junk);
				throw;
			}
			FSX_HIP(hipStreamEndCapture(ctx->stream, &en.g));
			FSX_HIP(hipGraphInstantiate(&en.ge, en.g, nullptr, nullptr, 0));
			cache.mv[key] = en;
			ge = en.ge;
		}
		else
			ge = it->second.ge;
	}
	bcg_state * flags = pinned + 1;
	hipEvent_t ev[2];
	FSX_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
	FSX_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
	auto launch_chunk = [&](int slot) {
		if (ge) FSX_HIP(hipGraphLaunch(ge, ctx->stream));
		else
			for (int k = 0; k < K; ++k) bcg_enqueue_iteration<V, I>(ctx, B, st);
		ctx->launches += 10 * K;
		FSX_HIP(hipMemcpyAsync(&flags[slot], st, sizeof(bcg_state), hipMemcpyDeviceToHost,
		                       ctx->stream));
		FSX_HIP(hipEventRecord(ev[slot], ctx->stream));
	};
	const int remaining = st_host.maxiter - st_host.iter;
	const int max_chunks = (remaining + K - 1) / K;
	int launched = 0, checked = 0;
	if (max_chunks > 0) {
		launch_chunk(0);
		launched = 1;
	}
	while (checked < launched) {
		if (launched < max_chunks) {
			launch_chunk(launched & 1);
			++launched;
		}
		FSX_HIP(hipEventSynchronize(ev[checked & 1]));
		ctx->host_syncs++;
		const bcg_state & f = flags[checked & 1];
		++checked;
		if (f.done) break;
	}
	if (launched > 0 && B.n > 0) {
		hipLaunchKernelGGL(bcg_xearly_kernel, dim3((unsigned)((B.n + 255) / 256)), dim3(256),
		                   0, ctx->stream, B.n, B.x, (const double *)B.ph,
		                   (const bcg_state *)st);
		FSX_HIP(hipGetLastError());
	}
	FSX_HIP(hipMemcpyAsync(pinned, st, sizeof(bcg_state), hipMemcpyDeviceToHost,
	                       ctx->stream));
	FSX_HIP(hipStreamSynchronize(ctx->stream));
	(void)hipEventDestroy(ev[0]);
	(void)hipEventDestroy(ev[1]);
	std::memcpy(&st_host, pinned, sizeof(bcg_state));
}

// ---- fused GMRES (op::gmres::apply, solvers/gmres.hh:79-258) -------------
//
// Same Arnoldi sequence, modified Gram-Schmidt order, Givens rotations
// (Bindel et al.), restart rule, back_solve and correct as the reference,
// with the Hessenberg column, rotations and least-squares right-hand side
// on the device.  Step k of a restart cycle:
//   G0  v = A P b_k (right; P = D^-1 enters as a scaled gather
//       0 + dinv[c] b_k[c], bitwise the z the reference forms) or
//       v = P A b_k (left); partials of v.b_0
//   F0  a one-block kernel stores h_0k
//   then for j = 0..k:
//   Gj  v = v - h_jk b_j; partials of v.b_{j+1} (j < k) or |v|^2 (j = k);
//       for j > 0 each block first sums G(j-1)'s partials into h_jk itself
//   GN  h_{k+1,k} = |v| from Gk's partials (each block); b_{k+1} = v / |v|
//       (v if |v| = 0)
//   GR  one thread: previous rotations on column k, the new rotation, the
//       rhs update, |dw_{k+1}| test, iteration count, end of cycle.
// A cycle's m steps are captured in one graph; the host runs back_solve,
// correct and the restart residual between cycles (every m iterations).

struct gm_state {
	double tol, vnorm;
	int iter, k, m, maxiter, done, status, iters, cycle_end;
};

struct gather_dinv {
	using value_type = double;
	const double * dinv;
	const double * x;
	__device__ __forceinline__ void init() {}
	__device__ __forceinline__ double operator()(int64_t c) const { return 0.0 + dinv[c] * x[c]; }
	__device__ __forceinline__ double at32(uint32_t c) const {
		return 0.0 + ld32(dinv, c) * ld32(x, c);
	}
};

template<class ACC, class T, bool LEFTJ>
struct epi_gm_v {
	static constexpr int NV = 1;
	static constexpr bool reduces = true;
	T * v;
	const T * b0;
	const T * dinv;
	double * partials;
	const gm_state * st;
	__device__ bool skip() const { return st->done != 0; }
	template<class XI>
	__device__ void row(int64_t i, ACC s, double (&d)[NV], XI &&) const {
		const double b0i = (double)b0[i]; // load before the store
		T vi = (T)s;
		if constexpr (LEFTJ) vi = 0.0 + dinv[i] * vi; // P.apply(z, v)
		v[i] = vi;
		d[0] += (double)vi * b0i; // v.dot(basis[0])
	}
	__device__ void block_end(const double (&d)[NV], unsigned blk, unsigned) const {
		store_block_partials<NV>(d, partials, blk);
	}
};

// H(row, col) = sum (or its square root for the norm), column-major with
// leading dimension ld (mdcolex of the reference)
struct fin_gm_h {
	const gm_state * st;
	double * H;
	int idx;
	bool norm;
	__device__ bool skip() const { return st->done != 0; }
	__device__ void operator()(const double (&t)[1]) const {
		H[idx] = norm ? sqrt(t[0]) : t[0];
	}
};

// rows per thread of the Gram-Schmidt passes (fewer blocks: fewer partials
// for every block of the next pass to sum)
constexpr int kGmPerThread = 8;

// v = v - h_jk b_j; partials of v.b_next (or |v|^2 when b_next is null)
// into `partials`.  h_jk is *hjk (j = 0: stored by the product's finalize)
// or, when `prev` is given, the sum of the previous pass's nprev partials,
// which every block forms in the same order while its vector loads are in
// flight (block 0 stores it at *hjk for the rotations): no finalize launch
// between passes.
__global__ __launch_bounds__(kUpdThreads) void gm_mgs_kernel(
	size_t n, double * __restrict__ v, const double * __restrict__ bj,
	const double * __restrict__ bnext, const double * __restrict__ prev, unsigned nprev,
	double * hjk, double * __restrict__ partials, const gm_state * st) {
	if (st->done) return;
	const unsigned blk = blockIdx.x;
	const size_t base = (size_t)blk * kUpdThreads * kGmPerThread + threadIdx.x;
	double bv[kGmPerThread], vv[kGmPerThread], nv[kGmPerThread];
#pragma unroll
	for (int j = 0; j < kGmPerThread; ++j) {
		const size_t i = base + (size_t)j * kUpdThreads;
		if (i < n) {
			bv[j] = bj[i];
			vv[j] = v[i];
			nv[j] = bnext ? bnext[i] : 0.0;
		}
	}
	double h;
	if (prev) {
		double t[1];
		sum_partials<1, kUpdThreads>(prev, nprev, t);
		h = t[0];
		if (blk == 0 && threadIdx.x == 0) *hjk = h;
	}
	else
		h = *hjk;
	double acc = 0.0;
#pragma unroll
	for (int j = 0; j < kGmPerThread; ++j) {
		const size_t i = base + (size_t)j * kUpdThreads;
		if (i < n) {
			const double vi = (-h) * bv[j] + vv[j]; // v.axpy(-h_jk, basis[j], v)
			v[i] = vi;
			acc += bnext ? vi * nv[j] : vi * vi;
		}
	}
	double d[1] = {acc};
	store_block_partials<1>(d, partials, blk);
}

// h_{k+1,k} = |v| from the last pass's nprev partials (every block, same
// order; block 0 stores it) and b_{k+1} = v scaled by 1/|v| (gmres.hh:140-146)
__global__ __launch_bounds__(kUpdThreads) void gm_scale_kernel(
	size_t n, const double * __restrict__ v, double * __restrict__ bnext,
	const double * __restrict__ prev, unsigned nprev, double * hnorm, const gm_state * st) {
	if (st->done) return;
	const size_t base = (size_t)blockIdx.x * kUpdThreads * kGmPerThread + threadIdx.x;
	double vv[kGmPerThread];
#pragma unroll
	for (int j = 0; j < kGmPerThread; ++j) {
		const size_t i = base + (size_t)j * kUpdThreads;
		if (i < n) vv[j] = v[i];
	}
	double t[1];
	sum_partials<1, kUpdThreads>(prev, nprev, t);
	const double vn = sqrt(t[0]);
	if (blockIdx.x == 0 && threadIdx.x == 0) *hnorm = vn;
#pragma unroll
	for (int j = 0; j < kGmPerThread; ++j) {
		const size_t i = base + (size_t)j * kUpdThreads;
		if (i < n) bnext[i] = vn != 0.0 ? (1.0 / vn) * vv[j] : vv[j];
	}
}

// Givens rotations of column k, |dw_{k+1}| test (gmres.hh:148-208, 282-331)
__global__ void gm_givens_kernel(double * H, int ld, double * cs, double * sn, double * dw,
                                 int k, gm_state * st) {
	if (st->done || threadIdx.x != 0) return;
	auto h = [&](int i, int j) -> double & { return H[i + (size_t)ld * j]; };
	const double vnorm0 = h(k + 1, k);
	for (int i = 0; i < k; ++i) {
		const double x = h(i, k), y = h(i + 1, k), c = cs[i], s = sn[i];
		h(i, k) = c * x + s * y;
		h(i + 1, k) = -s * x + c * y;
	}
	if (vnorm0 != 0.0) {
		const double f = h(k, k), g = h(k + 1, k);
		double c, s;
		if (g == 0.0) {
			c = 1.0;
			s = 0.0;
		}
		else if (f == 0.0) {
			c = 0.0;
			s = (g < 0.0) ? -1.0 : 1.0;
		}
		else {
			double r = sqrt(f * f + g * g);
			r = 1.0 / r;
			c = fabs(f) * r;
			s = copysign(g * r, f);
		}
		cs[k] = c;
		sn[k] = s;
		const double x = h(k, k), y = h(k + 1, k);
		h(k, k) = c * x + s * y;
		h(k + 1, k) = 0.0;
		(void)y;
		const double w = dw[k];
		dw[k] = c * w;
		dw[k + 1] = -s * w;
	}
	const double vnorm = fabs(dw[k + 1]);
	st->vnorm = vnorm;
	const int it = st->iter;
	st->iter = it + 1;
	st->k = k + 1;
	if (vnorm < st->tol) {
		st->status = FSX_CONVERGED_RTOL;
		st->iters = it + 1;
		st->done = 1;
		return;
	}
	if (it + 1 >= st->maxiter) {
		st->done = 1;
		return;
	}
	if (k + 1 == st->m) {
		st->cycle_end = 1;
		st->done = 1;
	}
}

// correct (gmres.hh:244-263): right: z = sum_i dy_i b_i in order, x += P z;
// left: x += dy_i b_i in order.  One pass over the basis per element.
template<bool RIGHT, bool JAC>
__global__ __launch_bounds__(256) void gm_correct_kernel(size_t n, double * __restrict__ x,
                                                         const double * const * basis,
                                                         const double * dy, int nr,
                                                         const double * __restrict__ dinv) {
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i >= n) return;
	if constexpr (RIGHT) {
		double z = 0.0;
		for (int j = 0; j <= nr; ++j) z = dy[j] * basis[j][i] + z;
		const double v = JAC ? 0.0 + dinv[i] * z : z;
		x[i] = 1.0 * v + x[i];
	}
	else {
		double xi = x[i];
		for (int j = 0; j <= nr; ++j) xi = dy[j] * basis[j][i] + xi;
		x[i] = xi;
	}
}

struct gm_bufs {
	const fsx_csr * A;
	double * v;
	const double * const * basis_dev; // device array of basis pointers
	std::vector<double *> basis;      // host copy
	const double * dinv;
	bool right;
	double *H, *cs, *sn, *dw;
	int ld, m;
	size_t n;
};

template<class V, class I>
static void gm_enqueue_cycle(fsx_ctx * ctx, const gm_bufs & B, gm_state * st) {
	hipStream_t s = ctx->stream;
	double * part = ctx->ws.partials;
	const size_t per = (size_t)kUpdThreads * kGmPerThread;
	const unsigned gu = (unsigned)((B.n + per - 1) / per);
	// the passes' partials alternate between two buffers, so that a pass
	// never overwrites the ones its blocks are still summing
	double * pp[2] = {part, part + gu};
	for (int k = 0; k < B.m; ++k) {
		const double * bk = B.basis[k];
		// G0: v = A P b_k / P A b_k, partials of v.b_0
		if (B.right && B.dinv) {
			epi_gm_v<double, double, false> e{B.v, B.basis[0], nullptr, part, st};
			spmv_launch_gather<decltype(e), double, V, I>(s, B.A, gather_dinv{B.dinv, bk}, e);
		}
		else if (!B.right && B.dinv) {
			epi_gm_v<double, double, true> e{B.v, B.basis[0], B.dinv, part, st};
			spmv_launch_typed<decltype(e), double, V, I, double>(s, B.A, bk, e);
		}
		else {
			epi_gm_v<double, double, false> e{B.v, B.basis[0], nullptr, part, st};
			spmv_launch_typed<decltype(e), double, V, I, double>(s, B.A, bk, e);
		}
		// h_0k from the product's partials
		spmv_finalize<1>(s, B.A, part, fin_gm_h{st, B.H, B.ld * k, false});
		for (int j = 0; j <= k; ++j)
			hipLaunchKernelGGL(gm_mgs_kernel, dim3(gu), dim3(kUpdThreads), 0, s, B.n, B.v,
			                   (const double *)B.basis[j],
			                   j < k ? (const double *)B.basis[j + 1] : nullptr,
			                   j > 0 ? (const double *)pp[(j - 1) & 1] : nullptr, gu,
			                   B.H + j + (size_t)B.ld * k, pp[j & 1], (const gm_state *)st);
		// GN: h_{k+1,k} = |v|; b_{k+1} = v / |v|
		hipLaunchKernelGGL(gm_scale_kernel, dim3(gu), dim3(kUpdThreads), 0, s, B.n,
		                   (const double *)B.v, B.basis[k + 1], (const double *)pp[k & 1], gu,
		                   B.H + k + 1 + (size_t)B.ld * k, (const gm_state *)st);
		// GR
		hipLaunchKernelGGL(gm_givens_kernel, dim3(1), dim3(64), 0, s, B.H, B.ld, B.cs, B.sn,
		                   B.dw, k, st);
	}
	FSX_HIP(hipGetLastError());
}

} // namespace fsx

// ---- scratch for the no-allocation CG entry point ---------------------------
//
// fsx_cg_solve_noalloc keeps everything a CG solve needs beyond its inputs and
// outputs -- the state, setup scalars, reduction partials and the work
// vectors r, z, p, w -- in one file-scope device array.  Each GPU holds its
// own copy (34 MiB) from the library's load until the process ends.  Host
// code never passes the array's name to a kernel: it is a host-side stand-in,
// not a device address.  fsx_ctx_create looks the address up on its device
// (hipGetSymbolAddress, once per device) and the entry point uses that.

namespace fsx {
constexpr size_t kNoallocMaxRows = FSX_NOALLOC_MAX_ROWS;
constexpr size_t kNoallocMaxPartials = (size_t)1 << 18; // doubles
constexpr size_t kNoallocVecBytes = kNoallocMaxRows * sizeof(double);
constexpr size_t kNoallocStateOff = 0;
constexpr size_t kNoallocSlotOff = 256;
constexpr size_t kNoallocPartOff = 4096;
constexpr size_t kNoallocVecOff = kNoallocPartOff + kNoallocMaxPartials * sizeof(double);
constexpr size_t kNoallocBytes = kNoallocVecOff + 4 * kNoallocVecBytes;
static_assert(sizeof(cg_state) <= kNoallocSlotOff &&
                  kNoallocSlotOff + 5 * sizeof(double) <= kNoallocPartOff,
              "scratch header layout");
static_assert(kNoallocVecOff % 256 == 0 && kNoallocVecBytes % 256 == 0, "alignment");
} // namespace fsx

static __device__ __align__(256) unsigned char scratch[fsx::kNoallocBytes];

namespace fsx {
constexpr int kNoallocMaxDevices = 64;
static std::mutex g_scratch_mu;
static void * g_scratch_ptr[kNoallocMaxDevices];
static hipError_t g_scratch_err[kNoallocMaxDevices];
static bool g_scratch_done[kNoallocMaxDevices];

void noalloc_init(int device) {
	if (device < 0 || device >= kNoallocMaxDevices) return;
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (g_scratch_done[device]) return;
	void * p = nullptr;
	const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
	g_scratch_err[device] = e;
	g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
	g_scratch_done[device] = true;
}

// the scratch array's device address on `device` (set up by fsx_ctx_create)
static char * noalloc_base(int device) {
	FSX_REQUIRE(device >= 0 && device < kNoallocMaxDevices, "cg_solve_noalloc: device index");
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (!g_scratch_done[device])
		throw error(FSX_ERR_STATE, "cg_solve_noalloc: scratch address not looked up on this "
		                           "device (no context was created on it)");
	if (g_scratch_err[device] != hipSuccess)
		throw error(FSX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
		                             hipGetErrorString(g_scratch_err[device]));
	return (char *)g_scratch_ptr[device];
}
} // namespace fsx

using namespace fsx;

namespace fsx {
// Average duration of `reps` back-to-back launches of the fused CG's K1 on A
// (NV = 1: partials of w.p; NV = 2: of w.p and w.w, the single-reduction
// inner CG), on the context stream.
template<int NV>
static void time_cg_spmv(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * p, fsx_vec * w,
                         int reps, double * avg_ms) {
	FSX_REQUIRE(ctx && A && avg_ms, "null argument");
	check_vec(p, "p");
	check_vec(w, "w");
	check_same(p, w);
	FSX_REQUIRE(A->nrows == A->ncols && (int64_t)p->n == A->nrows, "shape mismatch");
	FSX_REQUIRE(p->dt == FSX_F32 || A->vt == FSX_F64, "FP64 vectors need an FP64 matrix");
	FSX_REQUIRE(reps > 0, "reps must be > 0");
	// a state whose `done` flag stays clear (alpha etc. are not read by K1)
	dbuf stb(sizeof(cg_state));
	FSX_HIP(hipMemsetAsync(stb.p, 0, sizeof(cg_state), ctx->stream));
	double * part = ensure_partials(ctx, spmv_partials_size(A, NV));
	const cg_state * st = (const cg_state *)stb.p;
	auto launch = [&] {
		by_matrix_type(A, [&](auto * vtag, auto * itag) {
			using V = std::remove_pointer_t<decltype(vtag)>;
			using I = std::remove_pointer_t<decltype(itag)>;
			auto go = [&](auto * ttag) {
				using T = std::remove_pointer_t<decltype(ttag)>;
				using ACC = cg_acc_t<T, V>;
				if constexpr (NV == 1) {
					epi_cg_spmv<ACC, T> e{(T *)w->ptr(), part, st};
					spmv_launch_typed<decltype(e), ACC, V, I, T>(ctx->stream, A,
					                                             (const T *)p->ptr(), e);
				}
				else {
					epi_cg_spmv2<ACC, T> e{(T *)w->ptr(), part, st};
					spmv_launch_typed<decltype(e), ACC, V, I, T>(ctx->stream, A,
					                                             (const T *)p->ptr(), e);
				}
			};
			if (p->dt == FSX_F64) go((double *)nullptr);
			else go((float *)nullptr);
		});
	};
	launch(); // warm
	hipEvent_t e0, e1;
	FSX_HIP(hipEventCreate(&e0));
	FSX_HIP(hipEventCreate(&e1));
	FSX_HIP(hipEventRecord(e0, ctx->stream));
	for (int i = 0; i < reps; ++i) launch();
	FSX_HIP(hipEventRecord(e1, ctx->stream));
	FSX_HIP(hipEventSynchronize(e1));
	float ms = 0;
	FSX_HIP(hipEventElapsedTime(&ms, e0, e1));
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	*avg_ms = (double)ms / reps;
}
} // namespace fsx

extern "C" {

int fsx_cg_solve(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv,
                 const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                 fsx_info * info) {
	return guard([&] {
		FSX_REQUIRE(ctx && A && s && info, "null argument");
		check_vec(b, "b");
		check_vec(x, "x");
		check_same(b, x);
		FSX_REQUIRE(A->nrows == A->ncols && (int64_t)b->n == A->nrows,
		            "cg: shape mismatch");
		if (dinv) check_same(dinv, b);
		FSX_REQUIRE(b->dt == FSX_F32 || A->vt == FSX_F64,
		            "cg: FP64 vectors need an FP64 matrix");
		FSX_REQUIRE(b->ptr() != x->ptr() || b->n == 0, "cg: b and x must differ");
		const size_t n = b->n;
		const cg_work wk = pooled_work(ctx, A, n, b->dt, dinv != nullptr, 0, s->maxiter);
		int cnt = 0;
		if (b->dt == FSX_F64) cg_solve_impl<double>(ctx, A, dinv, b, x, *s, info, wk, &cnt);
		else cg_solve_impl<float>(ctx, A, dinv, b, x, *s, info, wk, &cnt);
		ctx->last_cg_hist.assign(cnt, 0.0);
		if (cnt > 0)
			FSX_HIP(hipMemcpy(ctx->last_cg_hist.data(), ctx->cg_cache->histbuf->p,
			                  sizeof(double) * cnt, hipMemcpyDeviceToHost));
	});
}

int fsx_cg_solve_diag(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv,
                      const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                      fsx_diag_fn diag, void * user, fsx_info * info) {
	return guard([&] {
		FSX_REQUIRE(ctx && A && s && info && diag, "null argument");
		check_vec(b, "b");
		check_vec(x, "x");
		check_same(b, x);
		FSX_REQUIRE(A->nrows == A->ncols && (int64_t)b->n == A->nrows,
		            "cg: shape mismatch");
		if (dinv) check_same(dinv, b);
		FSX_REQUIRE(b->dt == FSX_F32 || A->vt == FSX_F64,
		            "cg: FP64 vectors need an FP64 matrix");
		FSX_REQUIRE(b->ptr() != x->ptr() || b->n == 0, "cg: b and x must differ");
		const size_t n = b->n;
		const cg_work wk = pooled_work(ctx, A, n, b->dt, dinv != nullptr, 0, s->maxiter);
		int cnt = 0;
		if (b->dt == FSX_F64)
			cg_solve_impl<double>(ctx, A, dinv, b, x, *s, info, wk, &cnt, diag, user);
		else cg_solve_impl<float>(ctx, A, dinv, b, x, *s, info, wk, &cnt, diag, user);
		ctx->last_cg_hist.assign(cnt, 0.0);
		if (cnt > 0)
			FSX_HIP(hipMemcpy(ctx->last_cg_hist.data(), ctx->cg_cache->histbuf->p,
			                  sizeof(double) * cnt, hipMemcpyDeviceToHost));
	});
}

int fsx_cg_solve_noalloc(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv,
                         const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                         fsx_info * info) {
	return guard([&] {
		FSX_REQUIRE(ctx && A && s && info, "null argument");
		check_vec(b, "b");
		check_vec(x, "x");
		check_same(b, x);
		FSX_REQUIRE(A->nrows == A->ncols && (int64_t)b->n == A->nrows,
		            "cg_noalloc: shape mismatch");
		if (dinv) check_same(dinv, b);
		FSX_REQUIRE(b->dt == FSX_F32 || A->vt == FSX_F64,
		            "cg_noalloc: FP64 vectors need an FP64 matrix");
		FSX_REQUIRE(b->ptr() != x->ptr() || b->n == 0, "cg_noalloc: b and x must differ");
		const size_t n = b->n;
		FSX_REQUIRE(n <= kNoallocMaxRows,
		            "cg_noalloc: " + std::to_string(n) + " rows exceed the scratch limit of " +
		                std::to_string(kNoallocMaxRows) + " (FSX_NOALLOC_MAX_ROWS)");
		FSX_REQUIRE(cg_partials_need(A, n) <= kNoallocMaxPartials,
		            "cg_noalloc: the matrix's reducing SpMV leaves " +
		                std::to_string(cg_partials_need(A, n)) +
		                " partials, more than the scratch array's " +
		                std::to_string(kNoallocMaxPartials));
		char * base = noalloc_base(ctx->device);
		char * vec = base + kNoallocVecOff;
		cg_work wk{};
		wk.st = (cg_state *)(base + kNoallocStateOff);
		wk.slots = (double *)(base + kNoallocSlotOff);
		wk.part = (double *)(base + kNoallocPartOff);
		wk.r = vec;
		wk.z = dinv ? vec + kNoallocVecBytes : nullptr;
		wk.p = vec + 2 * kNoallocVecBytes;
		wk.w = vec + 3 * kNoallocVecBytes;
		wk.hist = nullptr;
		wk.graphs = false;
		if (b->dt == FSX_F64) cg_solve_impl<double>(ctx, A, dinv, b, x, *s, info, wk, nullptr);
		else cg_solve_impl<float>(ctx, A, dinv, b, x, *s, info, wk, nullptr);
	});
}

int fsx_time_cg_spmv(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * p, fsx_vec * w,
                     int reps, double * avg_ms) {
	return guard([&] { time_cg_spmv<1>(ctx, A, p, w, reps, avg_ms); });
}

int fsx_time_cg_spmv_sr(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * p, fsx_vec * w,
                        int reps, double * avg_ms) {
	return guard([&] { time_cg_spmv<2>(ctx, A, p, w, reps, avg_ms); });
}

int fsx_cg_history(fsx_ctx * ctx, double * hist, int cap, int * count) {
	return guard([&] {
		FSX_REQUIRE(ctx && count, "null argument");
		int c = (int)ctx->last_cg_hist.size();
		*count = c;
		if (hist)
			for (int i = 0; i < std::min(c, cap); ++i) hist[i] = ctx->last_cg_hist[i];
	});
}

int fsx_ir_cg_solve(fsx_ctx * ctx, const fsx_csr * A64, const fsx_csr * A32,
                    int jacobi, const fsx_vec * b, fsx_vec * x,
                    double rtol_outer, int max_outer,
                    const fsx_settings * inner, fsx_ir_info * info) {
	return guard([&] {
		FSX_REQUIRE(ctx && A64 && A32 && inner && info, "null argument");
		check_vec(b, "b");
		check_vec(x, "x");
		check_same(b, x);
		FSX_REQUIRE(b->dt == FSX_F64, "IR: b and x are FP64");
		FSX_REQUIRE(A64->vt == FSX_F64 && A32->vt == FSX_F32,
		            "IR: needs FP64 and FP32 matrices");
		FSX_REQUIRE(A64->nrows == A64->ncols && A32->nrows == A64->nrows &&
		                A32->ncols == A64->ncols && (int64_t)b->n == A64->nrows,
		            "IR: shape mismatch");
		FSX_REQUIRE(b->ptr() != x->ptr() || b->n == 0, "IR: b and x must differ");
		FSX_REQUIRE(jacobi >= 0 && jacobi <= 2, "IR: jacobi must be 0, 1 or 2");
		FSX_REQUIRE(jacobi != 2 || A32->idx == A64->idx,
		            "IR: jacobi 2 needs A32 = fsx_csr_convert(A64, FSX_F32) holding "
		            "fsx_csr_set_values_f32_sym_scaled values");
		std::memset(info, 0, sizeof(*info));
		info->status = FSX_UNKNOWN;
		const size_t n = b->n;
		fsx_vec * r = pooled(ctx, n, FSX_F64, 10);
		fsx_vec * r32 = pooled(ctx, n, FSX_F32, 11);
		fsx_vec * d32 = pooled(ctx, n, FSX_F32, 12);
		fsx_vec * dinv32 = jacobi == 1 ? pooled(ctx, n, FSX_F32, 13) : nullptr;
		fsx_vec * sc = jacobi == 2 ? pooled(ctx, n, FSX_F64, 19) : nullptr;
		if (jacobi == 1) {
			int rc = fsx_csr_diag_inv(ctx, A64, dinv32);
			if (rc) throw error(rc, get_error());
		}
		if (jacobi == 2 && n) {
			// s = D^{-1/2} of A64 (the scaling A32's values were written with)
			int * flag = (int *)(ctx->ws.scalars + 9);
			FSX_HIP(hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
			sym_scale_vector(ctx, A64, (double *)sc->ptr(), flag);
			int bad = 0;
			FSX_HIP(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
			FSX_HIP(hipStreamSynchronize(ctx->stream));
			ctx->host_syncs++;
			FSX_REQUIRE(!bad, "IR: a diagonal entry of A64 is missing or not positive");
		}
		const double b_norm2 = reduce_host(ctx, RED_L2SQ, b, nullptr);
		double b_norm = std::sqrt(b_norm2);
		info->rhs_norm = b_norm;
		if (b_norm == 0.0) b_norm = 1.0;
		double * rn2 = ctx->ws.scalars + 8; // |r|^2 on the device
		// r = b - A64 x and |r|^2 in one pass (FP64 values, FP64 vectors)
		auto residual = [&]() {
			by_matrix_type(A64, [&](auto * vtag, auto * itag) {
				using V = std::remove_pointer_t<decltype(vtag)>;
				using I = std::remove_pointer_t<decltype(itag)>;
				epi_res_norm<double, double> e{(double *)r->ptr(),
				                               (const double *)b->ptr(),
				                               ctx->ws.partials};
				spmv_launch_typed<decltype(e), double, V, I, double>(
					ctx->stream, A64, (const double *)x->ptr(), e);
			});
			spmv_finalize<1>(ctx->stream, A64, ctx->ws.partials, fin_store<1>{rn2});
			FSX_HIP(hipGetLastError());
			ctx->launches += 2;
			return std::sqrt(fetch_scalar(ctx, rn2));
		};
		// partials for the outer residual and the inner solves, grown once
		// before any buffer address is taken
		ensure_partials(ctx, std::max(spmv_partials_size(A64, 1),
		                              cg_partials_need(A32, n, jacobi != 1)));
		fsx_settings is = *inner;
		is.use_zero_guess = 1;
		// inner work: pooled slots 14-17
		// unpreconditioned inner solves (jacobi 0, 2) take the single-reduction
		// iteration
		const cg_work iwk =
			pooled_work(ctx, A32, n, FSX_F32, jacobi == 1, 14, is.maxiter, jacobi != 1);
		double rnorm = residual();
		const unsigned g = (unsigned)((n + 255) / 256);
		int outer = 0;
		for (;;) {
			if (rnorm / b_norm <= rtol_outer) {
				info->status = FSX_CONVERGED_RTOL;
				break;
			}
			if (outer >= max_outer || rnorm == 0.0) {
				info->status = rnorm == 0.0 ? FSX_CONVERGED_RTOL : FSX_DIVERGED_ITERS;
				break;
			}
			// normalised residual in FP32 (scaled by s for the symmetrically
			// scaled inner matrix)
			if (sc)
				hipLaunchKernelGGL(ir_scale_sym_kernel, dim3(g), dim3(256), 0, ctx->stream, n,
				                   (const double *)r->ptr(), (const double *)sc->ptr(),
				                   (float *)r32->ptr(), (const double *)rn2);
			else
				hipLaunchKernelGGL(ir_scale_kernel, dim3(g), dim3(256), 0, ctx->stream,
				                   n, (const double *)r->ptr(), (float *)r32->ptr(),
				                   (const double *)rn2);
			FSX_HIP(hipGetLastError());
			fsx_info ii;
			int cnt = 0;
			cg_solve_impl<float>(ctx, A32, dinv32, r32, d32, is, &ii, iwk, &cnt);
			info->inner_iters_total += cnt;
			// x += |r| d  (FP64; d = s y for the scaled inner system)
			if (sc)
				hipLaunchKernelGGL(ir_update_sym_kernel, dim3(g), dim3(256), 0, ctx->stream, n,
				                   (double *)x->ptr(), (const float *)d32->ptr(),
				                   (const double *)sc->ptr(), (const double *)rn2);
			else
				hipLaunchKernelGGL(ir_update_kernel, dim3(g), dim3(256), 0, ctx->stream,
				                   n, (double *)x->ptr(), (const float *)d32->ptr(),
				                   (const double *)rn2);
			FSX_HIP(hipGetLastError());
			ctx->launches += 2;
			++outer;
			rnorm = residual();
		}
		info->outer_iters = outer;
		info->rel_res_final = rnorm / b_norm;
	});
}

int fsx_ir_cg_rr_solve(fsx_ctx * ctx, const fsx_csr * A64, const fsx_csr * A32,
                       const fsx_vec * b, fsx_vec * x, double rtol, int max_segments,
                       const fsx_settings * inner, fsx_ir_info * info) {
	return guard([&] {
		FSX_REQUIRE(ctx && A64 && A32 && inner && info, "null argument");
		check_vec(b, "b");
		check_vec(x, "x");
		check_same(b, x);
		FSX_REQUIRE(b->dt == FSX_F64, "RR: b and x are FP64");
		FSX_REQUIRE(A64->vt == FSX_F64 && A32->vt == FSX_F32,
		            "RR: needs FP64 and FP32 matrices");
		FSX_REQUIRE(A64->nrows == A64->ncols && A32->nrows == A64->nrows &&
		                A32->ncols == A64->ncols && (int64_t)b->n == A64->nrows,
		            "RR: shape mismatch");
		FSX_REQUIRE(b->ptr() != x->ptr() || b->n == 0, "RR: b and x must differ");
		FSX_REQUIRE(A32->idx == A64->idx,
		            "RR: needs A32 = fsx_csr_convert(A64, FSX_F32) holding "
		            "fsx_csr_set_values_f32_sym_scaled values");
		FSX_REQUIRE(inner->rtol > 0.0 && inner->rtol < 1.0,
		            "RR: the segment reduction inner->rtol must lie in (0, 1)");
		FSX_REQUIRE(inner->maxiter >= 1, "RR: inner->maxiter must be at least 1");
		FSX_REQUIRE(max_segments >= 0, "RR: max_segments must not be negative");
		std::memset(info, 0, sizeof(*info));
		info->status = FSX_UNKNOWN;
		const size_t n = b->n;
		fsx_vec * r = pooled(ctx, n, FSX_F64, 10);
		fsx_vec * y = pooled(ctx, n, FSX_F32, 12);
		fsx_vec * sc = pooled(ctx, n, FSX_F64, 19);
		if (n) {
			// s = D^{-1/2} of A64 (the scaling A32's values were written with)
			int * flag = (int *)(ctx->ws.scalars + 9);
			FSX_HIP(hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
			sym_scale_vector(ctx, A64, (double *)sc->ptr(), flag);
			int bad = 0;
			FSX_HIP(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
			FSX_HIP(hipStreamSynchronize(ctx->stream));
			ctx->host_syncs++;
			FSX_REQUIRE(!bad, "RR: a diagonal entry of A64 is missing or not positive");
		}
		const double b_norm2 = reduce_host(ctx, RED_L2SQ, b, nullptr);
		double b_norm = std::sqrt(b_norm2);
		info->rhs_norm = b_norm;
		if (b_norm == 0.0) b_norm = 1.0;
		double * rn2 = ctx->ws.scalars + 8;
		// partials for the FP64 residual and the segments, grown once before
		// any buffer address is taken
		ensure_partials(ctx, std::max(spmv_partials_size(A64, 1), cg_partials_need(A32, n, true)));
		const cg_work iwk = pooled_work(ctx, A32, n, FSX_F32, false, 14, inner->maxiter, true);
		auto residual = [&]() {
			by_matrix_type(A64, [&](auto * vtag, auto * itag) {
				using V = std::remove_pointer_t<decltype(vtag)>;
				using I = std::remove_pointer_t<decltype(itag)>;
				epi_res_norm<double, double> e{(double *)r->ptr(), (const double *)b->ptr(),
				                               ctx->ws.partials};
				spmv_launch_typed<decltype(e), double, V, I, double>(
					ctx->stream, A64, (const double *)x->ptr(), e);
			});
			spmv_finalize<1>(ctx->stream, A64, ctx->ws.partials, fin_store<1>{rn2});
			FSX_HIP(hipGetLastError());
			ctx->launches += 2;
			return std::sqrt(fetch_scalar(ctx, rn2));
		};
		const size_t per = (size_t)kUpdThreads * kUpdPerThread;
		const unsigned g2 = (unsigned)((n + per - 1) / per);
		const unsigned g = (unsigned)((n + 255) / 256);
		double * rho_slot = iwk.slots;
		// rh <- (float)(S r / c) (FIRST: p = rh) and |rh|^2
		auto replace = [&](bool first, double c) {
			if (first)
				hipLaunchKernelGGL((rr_replace_kernel<true>), dim3(g2), dim3(kUpdThreads), 0,
				                   ctx->stream, n, (const double *)r->ptr(),
				                   (const double *)sc->ptr(), 1.0 / c, (float *)iwk.r,
				                   (float *)iwk.p, (float *)y->ptr(), iwk.part);
			else
				hipLaunchKernelGGL((rr_replace_kernel<false>), dim3(g2), dim3(kUpdThreads), 0,
				                   ctx->stream, n, (const double *)r->ptr(),
				                   (const double *)sc->ptr(), 1.0 / c, (float *)iwk.r,
				                   (float *)iwk.p, (float *)y->ptr(), iwk.part);
			hipLaunchKernelGGL((finalize_kernel<1, fin_store<1>>), dim3(1), dim3(kFinThreads), 0,
			                   ctx->stream, (const double *)iwk.part, g2,
			                   fin_store<1>{rho_slot});
			FSX_HIP(hipGetLastError());
			ctx->launches += 2;
			return fetch_scalar(ctx, rho_slot);
		};
		double rnorm = n ? residual() : 0.0;
		int seg = 0;
		if (rnorm / b_norm <= rtol || rnorm == 0.0)
			info->status = FSX_CONVERGED_RTOL;
		else {
			const double c = rnorm;
			double rho = replace(true, c);
			cg_bufs B{A32, y->ptr(), iwk.r, iwk.r, iwk.p, iwk.w, nullptr, iwk.part,
			          iwk.part + spmv_partials_size(A32, 2), n};
			for (;;) {
				if (seg >= max_segments) {
					info->status = FSX_DIVERGED_ITERS;
					break;
				}
				if (!(rho > 0.0)) {
					// the scaled residual vanished in FP32 while the FP64 one
					// did not: no further progress is possible
					info->status = FSX_DIVERGED_BREAKDOWN;
					break;
				}
				// the segment ends after a reduction of inner->rtol, or once
				// the true residual is estimated below half the target (the
				// ratio |r| / |rh| of the last replacement)
				const double rh_norm = std::sqrt(rho);
				cg_state st{};
				st.rho1 = rho;
				st.rho0 = rho;
				st.res = rh_norm;
				st.tol = std::max(inner->rtol * rh_norm, 0.5 * rtol * b_norm * rh_norm / rnorm);
				st.maxiter = inner->maxiter;
				by_matrix_type(A32, [&](auto * vtag, auto * itag) {
					using V = std::remove_pointer_t<decltype(vtag)>;
					using I = std::remove_pointer_t<decltype(itag)>;
					run_cg_loop<float, V, I>(ctx, B, iwk, st, nullptr, nullptr, nullptr, nullptr);
				});
				info->inner_iters_total += st.iter;
				++seg;
				hipLaunchKernelGGL(rr_update_kernel, dim3(g), dim3(256), 0, ctx->stream, n,
				                   (double *)x->ptr(), (const float *)y->ptr(),
				                   (const double *)sc->ptr(), c);
				FSX_HIP(hipGetLastError());
				ctx->launches++;
				rnorm = residual();
				if (rnorm / b_norm <= rtol || rnorm == 0.0) {
					info->status = FSX_CONVERGED_RTOL;
					break;
				}
				rho = replace(false, c);
			}
		}
		info->outer_iters = seg;
		info->rel_res_final = rnorm / b_norm;
	});
}

int fsx_bicgstab_solve(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv,
                       const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                       fsx_info * info) {
	return guard([&] {
		FSX_REQUIRE(ctx && A && s && info, "null argument");
		check_vec(b, "b");
		check_vec(x, "x");
		check_same(b, x);
		FSX_REQUIRE(A->nrows == A->ncols && (int64_t)b->n == A->nrows,
		            "bicgstab: shape mismatch");
		FSX_REQUIRE(A->vt == FSX_F64 && b->dt == FSX_F64, "bicgstab: FP64 matrix and vectors");
		if (dinv) check_same(dinv, b);
		FSX_REQUIRE(b->ptr() != x->ptr() || b->n == 0, "bicgstab: b and x must differ");
		std::memset(info, 0, sizeof(*info));
		info->status = FSX_UNKNOWN;
		const size_t n = b->n;
		fsx_vec * res = pooled(ctx, n, FSX_F64, 20);
		fsx_vec * rt = pooled(ctx, n, FSX_F64, 21);
		fsx_vec * p = pooled(ctx, n, FSX_F64, 22);
		fsx_vec * v = pooled(ctx, n, FSX_F64, 23);
		fsx_vec * ph = dinv ? pooled(ctx, n, FSX_F64, 24) : p;
		fsx_vec * sv = pooled(ctx, n, FSX_F64, 25);
		fsx_vec * sh = dinv ? pooled(ctx, n, FSX_F64, 26) : sv;
		fsx_vec * t = pooled(ctx, n, FSX_F64, 27);
		// bicgstab.hh:52-79
		double b_norm = std::sqrt(reduce_host(ctx, RED_L2SQ, b, nullptr));
		if (b_norm == 0.0) b_norm = 1.0;
		const double tol = s->rtol * b_norm;
		info->rhs_norm = (float)b_norm;
		if (s->use_zero_guess) {
			info->sol_norm_initial = 0;
			launch_copy(ctx, b, res);
			launch_set(ctx, 0.0, x);
		}
		else {
			info->sol_norm_initial = (float)std::sqrt(reduce_host(ctx, RED_L2SQ, x, nullptr));
			int rc = fsx_csr_residual(ctx, A, b, x, res);
			if (rc) throw error(rc, get_error());
		}
		double res_norm = std::sqrt(reduce_host(ctx, RED_L2SQ, res, nullptr));
		double rt_norm = res_norm;
		info->res_norm_initial = (float)res_norm;
		if (res_norm < tol) {
			info->status = FSX_CONVERGED_RTOL;
			info->res_norm_final = (float)res_norm;
			return;
		}
		// alpha = 1, omega = 1, rho = {2, 1} (list-initialised, bicgstab.hh:85-89)
		bcg_state st{};
		st.alpha = 1.0;
		st.omega = 1.0;
		st.beta = 0.0;
		st.rho0 = 2.0;
		st.rho1 = 1.0;
		st.tol = tol;
		st.rtol_abs = (double)s->rtol;
		st.maxiter = s->maxiter;
		st.res = res_norm;
		launch_copy(ctx, res, rt);
		launch_set(ctx, 0.0, p);
		launch_set(ctx, 0.0, v);
		bcg_bufs B{A, (double *)x->ptr(), (double *)res->ptr(), (double *)rt->ptr(),
		           (double *)p->ptr(), (double *)v->ptr(), (double *)ph->ptr(),
		           (double *)sv->ptr(), (double *)sh->ptr(), (double *)t->ptr(),
		           dinv ? (const double *)dinv->ptr() : nullptr, n};
		int it = 0;
		bool finished = false;
		while (!finished) {
			// top of iteration `it` on the host: rho, breakdown restart
			// (bicgstab.hh:105-116), beta (118)
			bool run = false;
			while (it < s->maxiter) {
				const double rho = reduce_host(ctx, RED_DOT, rt, res);
				if (std::sqrt(std::fabs(rho)) < DBL_EPSILON * rt_norm) {
					int rc = fsx_csr_residual(ctx, A, b, x, res);
					if (rc) throw error(rc, get_error());
					launch_copy(ctx, res, rt);
					res_norm = std::sqrt(reduce_host(ctx, RED_L2SQ, res, nullptr));
					rt_norm = res_norm;
					launch_copy(ctx, res, p);
					++info->restarts;
					++it;
					continue;
				}
				st.rho1 = rho;
				if (it > 0) st.beta = (rho / st.rho0) * (st.alpha / st.omega);
				run = true;
				break;
			}
			if (!run) break; // maxiter reached in the restart loop
			st.iter = it;
			st.first = it == 0;
			st.rt_norm = rt_norm;
			st.done = st.restart = st.early = 0;
			st.status = FSX_UNKNOWN;
			st.iters = 0;
			by_matrix_type(A, [&](auto * vtag, auto * itag) {
				using V = std::remove_pointer_t<decltype(vtag)>;
				using I = std::remove_pointer_t<decltype(itag)>;
				bcg_run_chunks<V, I>(ctx, B, st);
			});
			res_norm = st.res;
			if (st.restart) {
				// the device found r~ orthogonal to res at the top of st.iter
				int rc = fsx_csr_residual(ctx, A, b, x, res);
				if (rc) throw error(rc, get_error());
				launch_copy(ctx, res, rt);
				res_norm = std::sqrt(reduce_host(ctx, RED_L2SQ, res, nullptr));
				rt_norm = res_norm;
				launch_copy(ctx, res, p);
				++info->restarts;
				it = st.iter + 1;
				continue;
			}
			finished = true;
		}
		info->iters = st.iters;
		if (st.iters || st.status == FSX_DIVERGED_BREAKDOWN) info->status = st.status;
		info->res_norm_final = (float)res_norm;
		info->sol_norm_final = (float)std::sqrt(reduce_host(ctx, RED_L2SQ, x, nullptr));
		if (info->iters == 0) info->status = FSX_DIVERGED_ITERS;
	});
}

int fsx_gmres_solve(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv,
                    const fsx_vec * b, fsx_vec * x, const fsx_settings * s_in,
                    fsx_info * info) {
	return guard([&] {
		FSX_REQUIRE(ctx && A && s_in && info, "null argument");
		check_vec(b, "b");
		check_vec(x, "x");
		check_same(b, x);
		FSX_REQUIRE(A->nrows == A->ncols && (int64_t)b->n == A->nrows,
		            "gmres: shape mismatch");
		FSX_REQUIRE(A->vt == FSX_F64 && b->dt == FSX_F64, "gmres: FP64 matrix and vectors");
		if (dinv) check_same(dinv, b);
		FSX_REQUIRE(b->ptr() != x->ptr() || b->n == 0, "gmres: b and x must differ");
		// settings::validate (gmres.hh:394-405)
		fsx_settings s = *s_in;
		if (s.max_krylov_dim < 0) s.max_krylov_dim = s.maxiter;
		FSX_REQUIRE(s.max_krylov_dim <= 100, "GMRES: max_krylov_dim is larger than bound");
		if (!s.restart)
			FSX_REQUIRE(s.maxiter <= s.max_krylov_dim,
			            "GMRES: maxiters must be less than or equal to max_krylov_dim when "
			            "not using restart");
		std::memset(info, 0, sizeof(*info));
		info->status = FSX_UNKNOWN;
		const size_t n = b->n;
		const int max_dim = std::max(0, std::min(s.max_krylov_dim, s.maxiter));
		const int ld = max_dim + 1;
		const bool right = s.pre_side != 0;
		// pooled vectors: res, z, v, basis[0..max_dim]; small device arrays
		fsx_vec * res = pooled(ctx, n, FSX_F64, 30);
		fsx_vec * z = pooled(ctx, n, FSX_F64, 31);
		fsx_vec * v = pooled(ctx, n, FSX_F64, 32);
		std::vector<fsx_vec *> bvec(ld);
		std::vector<double *> basis(ld);
		for (int i = 0; i < ld; ++i) {
			bvec[i] = pooled(ctx, n, FSX_F64, 40 + i);
			basis[i] = (double *)bvec[i]->ptr();
		}
		const size_t small = (size_t)ld * ld + 4 * (size_t)ld;
		fsx_vec * smallv = pooled(ctx, small + (size_t)ld, FSX_F64, 39); // + basis ptrs
		double * H = (double *)smallv->ptr();
		double * cs = H + (size_t)ld * ld;
		double * sn = cs + ld;
		double * dw = sn + ld;
		double * dy = dw + ld;
		double ** bptr = (double **)(dy + ld);
		// gmres reset(): H, cos, sin, dw, dy zero
		FSX_HIP(hipMemsetAsync(H, 0, small * sizeof(double), ctx->stream));
		FSX_HIP(hipMemcpyAsync(bptr, basis.data(), ld * sizeof(double *),
		                       hipMemcpyHostToDevice, ctx->stream));
		double b_norm = std::sqrt(reduce_host(ctx, RED_L2SQ, b, nullptr));
		info->rhs_norm = (float)b_norm;
		if (b_norm < DBL_EPSILON) b_norm = 1.0;
		const double tol = s.rtol * b_norm;
		if (s.use_zero_guess) launch_set(ctx, 0.0, x);
		auto precond = [&](const fsx_vec * in, fsx_vec * out) {
			if (dinv) {
				int rc = fsx_jacobi_apply(ctx, dinv, in, out);
				if (rc) throw error(rc, get_error());
			}
			else
				launch_copy(ctx, in, out);
		};
		auto residual_into = [&](fsx_vec * out) {
			int rc = fsx_csr_residual(ctx, A, b, x, out);
			if (rc) throw error(rc, get_error());
		};
		fsx_vec * b0 = bvec[0];
		if (!right) {
			if (s.use_zero_guess) launch_copy(ctx, b, b0);
			else residual_into(b0);
			precond(b0, res);
		}
		else {
			if (s.use_zero_guess) launch_copy(ctx, b, res);
			else residual_into(res);
		}
		const double beta = std::sqrt(reduce_host(ctx, RED_L2SQ, res, nullptr));
		info->res_norm_initial = (float)beta;
		if (beta < tol) {
			info->res_norm_final = (float)beta;
			info->status = FSX_CONVERGED_RTOL;
			return;
		}
		// res.scale(1/beta); basis[0].copy(res); dwvec[0] = beta
		auto start_cycle = [&](double bt) {
			fsx_vec * tmp = nullptr;
			(void)tmp;
			int rc = fsx_scale(ctx, 1.0 / bt, res);
			if (rc) throw error(rc, get_error());
			launch_copy(ctx, res, b0);
			FSX_HIP(hipMemcpyAsync(dw, &bt, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
			FSX_HIP(hipStreamSynchronize(ctx->stream)); // bt is a host local
		};
		start_cycle(beta);
		gm_bufs B{A, (double *)v->ptr(), bptr, basis, dinv ? (const double *)dinv->ptr() : nullptr,
		          right, H, cs, sn, dw, ld, max_dim, n};
		auto & cache = cache_of(ctx);
		gm_state * st = (gm_state *)cache.statebuf->p;
		gm_state * pinned = (gm_state *)ctx->host_scalars;
		static_assert(sizeof(gm_state) <= sizeof(cg_state) * 4, "");
		const size_t per = (size_t)kUpdThreads * kGmPerThread;
		ensure_partials(ctx, std::max<size_t>(spmv_partials_size(A, 1), 2 * ((n + per - 1) / per)));
		gm_state sh{};
		sh.tol = tol;
		sh.m = max_dim;
		sh.maxiter = s.maxiter;
		sh.vnorm = beta;
		std::vector<double> hH, hdw, hdy;
		auto back_solve_correct = [&](int nr) {
			// back_solve (gmres.hh:333-348) on the host, correct on the device
			hH.resize((size_t)ld * ld);
			hdw.resize(ld);
			hdy.assign(ld, 0.0);
			FSX_HIP(hipMemcpyAsync(hH.data(), H, hH.size() * sizeof(double),
			                       hipMemcpyDeviceToHost, ctx->stream));
			FSX_HIP(hipMemcpyAsync(hdw.data(), dw, ld * sizeof(double), hipMemcpyDeviceToHost,
			                       ctx->stream));
			FSX_HIP(hipStreamSynchronize(ctx->stream));
			auto h = [&](int i, int j) { return hH[i + (size_t)ld * j]; };
			hdy[nr] = hdw[nr] / h(nr, nr);
			for (int kk = nr - 1; kk >= 0; kk--) {
				hdy[kk] = hdw[kk];
				for (int i = kk + 1; i <= nr; i++) hdy[kk] -= h(kk, i) * hdy[i];
				hdy[kk] = hdy[kk] / h(kk, kk);
			}
			FSX_HIP(hipMemcpyAsync(dy, hdy.data(), ld * sizeof(double), hipMemcpyHostToDevice,
			                       ctx->stream));
			const unsigned g1 = (unsigned)((n + 255) / 256);
			if (right && dinv)
				hipLaunchKernelGGL((gm_correct_kernel<true, true>), dim3(g1), dim3(256), 0,
				                   ctx->stream, n, (double *)x->ptr(),
				                   (const double * const *)bptr, (const double *)dy, nr,
				                   (const double *)dinv->ptr());
			else if (right)
				hipLaunchKernelGGL((gm_correct_kernel<true, false>), dim3(g1), dim3(256), 0,
				                   ctx->stream, n, (double *)x->ptr(),
				                   (const double * const *)bptr, (const double *)dy, nr,
				                   (const double *)nullptr);
			else
				hipLaunchKernelGGL((gm_correct_kernel<false, false>), dim3(g1), dim3(256), 0,
				                   ctx->stream, n, (double *)x->ptr(),
				                   (const double * const *)bptr, (const double *)dy, nr,
				                   (const double *)nullptr);
			FSX_HIP(hipGetLastError());
			FSX_HIP(hipStreamSynchronize(ctx->stream)); // hdy is reused
		};
		int k = 0;
		while (max_dim > 0) {
			sh.k = 0;
			sh.done = sh.cycle_end = 0;
			std::memcpy(pinned, &sh, sizeof(gm_state));
			FSX_HIP(hipMemcpyAsync(st, pinned, sizeof(gm_state), hipMemcpyHostToDevice,
			                       ctx->stream));
			by_matrix_type(A, [&](auto * vtag, auto * itag) {
				using V = std::remove_pointer_t<decltype(vtag)>;
				using I = std::remove_pointer_t<decltype(itag)>;
				hipGraphExec_t ge = nullptr;
				if (use_graphs()) {
					std::vector<const void *> key{(const void *)&gm_enqueue_cycle<V, I>,
					                              (const void *)B.v, (const void *)H,
					                              (const void *)B.dinv,
					                              (const void *)ctx->ws.partials,
					                              (const void *)st,
					                              (const void *)(intptr_t)n,
					                              (const void *)(intptr_t)max_dim,
					                              (const void *)(intptr_t)right};
					for (double * bp : basis) key.push_back(bp);
					append_layout(key, A);
					auto it = cache.mv.find(key);
					if (it == cache.mv.end()) {
						if (cache.mv.size() >= 32) cache.clear_graphs();
						cg_graph_cache::entry en;
						FSX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
						try {
							gm_enqueue_cycle<V, I>(ctx, B, st);
						}
						catch (...) {
							hipGraph_t junk = nullptr;
							(void)hipStreamEndCapture(ctx->stream, &junk);
							if (junk) (void)hipGraphDestroy(junk);
							throw;
						}
						FSX_HIP(hipStreamEndCapture(ctx->stream, &en.g));
						FSX_HIP(hipGraphInstantiate(&en.ge, en.g, nullptr, nullptr, 0));
						cache.mv[key] = en;
						ge = en.ge;
					}
					else
						ge = it->second.ge;
				}
				if (ge) FSX_HIP(hipGraphLaunch(ge, ctx->stream));
				else gm_enqueue_cycle<V, I>(ctx, B, st);
			});
			ctx->launches += (size_t)max_dim * (max_dim + 9) / 2;
			FSX_HIP(hipMemcpyAsync(pinned, st, sizeof(gm_state), hipMemcpyDeviceToHost,
			                       ctx->stream));
			FSX_HIP(hipStreamSynchronize(ctx->stream));
			ctx->host_syncs++;
			std::memcpy(&sh, pinned, sizeof(gm_state));
			k = sh.k;
			if (!sh.cycle_end) break; // converged or maxiter
			// restart (gmres.hh:215-233)
			back_solve_correct(k - 1);
			if (!right) {
				residual_into(b0);
				precond(b0, res);
			}
			else
				residual_into(res);
			const double betar = std::sqrt(reduce_host(ctx, RED_L2SQ, res, nullptr));
			start_cycle(betar);
			++info->restarts;
			k = 0;
		}
		if (k > 0) back_solve_correct(k - 1);
		info->iters = sh.iters;
		if (sh.iters) info->status = sh.status;
		info->res_norm_final = (float)sh.vnorm;
		info->sol_norm_final = (float)std::sqrt(reduce_host(ctx, RED_L2SQ, x, nullptr));
		if (info->iters == 0) info->status = FSX_DIVERGED_ITERS;
	});
}

int fsx_cg_update(fsx_ctx * ctx, double alpha, const fsx_vec * p, const fsx_vec * w,
                  const fsx_vec * dinv, fsx_vec * x, fsx_vec * r, fsx_vec * z, double * rr,
                  double * rz) {
	return guard([&] {
		FSX_REQUIRE(ctx, "null context");
		check_vec(p, "p");
		check_vec(w, "w");
		check_vec(x, "x");
		check_vec(r, "r");
		check_same(p, w);
		check_same(p, x);
		check_same(p, r);
		FSX_REQUIRE(p->dt == FSX_F64, "cg_update is FP64");
		if (dinv) {
			check_vec(z, "z");
			check_same(p, dinv);
			check_same(p, z);
		}
		const void * ptrs[] = {p->ptr(), w->ptr(), x->ptr(), r->ptr()};
		for (int a = 0; a < 4; ++a)
			for (int b2 = a + 1; b2 < 4; ++b2)
				FSX_REQUIRE(ptrs[a] != ptrs[b2] || p->n == 0, "cg_update: p, w, x, r must differ");
		const size_t n = p->n;
		const size_t per = (size_t)kUpdThreads * kUpdPerThread;
		const unsigned g = (unsigned)std::max<size_t>(1, (n + per - 1) / per);
		ensure_partials(ctx, 2 * (size_t)g);
		double * slot = ctx->ws.scalars + 2;
		if (dinv)
			hipLaunchKernelGGL((cg_update_full_kernel<true>), dim3(g), dim3(kUpdThreads), 0,
			                   ctx->stream, n, alpha, (double *)x->ptr(), (double *)r->ptr(),
			                   (const double *)p->ptr(), (const double *)w->ptr(),
			                   (const double *)dinv->ptr(), (double *)z->ptr(),
			                   ctx->ws.partials);
		else
			hipLaunchKernelGGL((cg_update_full_kernel<false>), dim3(g), dim3(kUpdThreads), 0,
			                   ctx->stream, n, alpha, (double *)x->ptr(), (double *)r->ptr(),
			                   (const double *)p->ptr(), (const double *)w->ptr(),
			                   (const double *)nullptr, (double *)nullptr, ctx->ws.partials);
		hipLaunchKernelGGL((finalize_kernel<2, fin_store<2>>), dim3(1), dim3(kFinThreads), 0,
		                   ctx->stream, (const double *)ctx->ws.partials, g,
		                   fin_store<2>{slot});
		FSX_HIP(hipGetLastError());
		ctx->launches += 2;
		double out[2];
		FSX_HIP(hipMemcpyAsync(ctx->host_scalars, slot, 2 * sizeof(double),
		                       hipMemcpyDeviceToHost, ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		out[0] = ctx->host_scalars[0];
		out[1] = ctx->host_scalars[1];
		if (rr) *rr = out[0];
		if (rz) *rz = out[1];
	});
}

} // extern "C"
