// fsx_amg.hip -- unsmoothed-aggregation AMG (mg::ua::bound_solver) on the
// device, for one colour.
//
// Reference: flecsolve/solvers/mg/ua.hh:94-200 (setup and apply),
// coarsen.hh:26-655 (pairwise aggregation and the Galerkin coarse
// operator), intergrid.hh:34-180 (prolong = injection, restrict = sum over
// the aggregate), cycle.hh:43-272 (V, relaxed W, relaxed kappa, K-cycles),
// cg_solve.hh (dense LU on the coarsest level), gs.hh (symmetric hybrid
// Gauss-Seidel smoothing on every level).
//
// Setup runs on the host, as in the reference (its coarsening tasks run on
// the CPU): pairwise aggregation, then the coarse matrix, level by level,
// until a level has fewer than min_coarse rows or max_levels is reached; the
// coarsest matrix is factored densely on the host.  Each level keeps its
// matrix as an fsx_csr (row sums in CSR order: the slice layouts, never the
// lanes-per-row mapping), a hybrid Gauss-Seidel smoother, the fine-to-coarse
// map and, per coarse row, its fine rows in ascending order.  A cycle runs on
// the device: smoothing sweeps, residuals and restriction/prolongation
// kernels; the coarsest solve goes through the host (a few hundred unknowns,
// solved serially as the reference does).  Each kernel repeats the
// reference's arithmetic in its order, so a V, relaxed W or relaxed kappa
// cycle is bit-identical to the sequential algorithm; the K-cycle's dot
// products are device reductions (rounding-level differences).
//
// Aggregation (coarsen.hh:44-392), restated with the reference's semantics:
//  * candidates (get_unmarked, 176-206): with the diagonal-dominance check,
//    row r stays out of the coarse levels when a_rr > 5 * offd_sum, where
//    offd_sum is an int that accumulates |a_rc| over the off-diagonal
//    entries with truncation after every addition (std::reduce with an int
//    initial value, then `+=` on that int);
//  * strength (soc_graph, 133-174): c is a strong neighbour of r when
//    c != r, c is a candidate (first pass only) and a_rc < beta * m_r, m_r
//    the smallest value stored in row r (diagonal included);
//  * priorities (prospect, 37-131): a candidate's priority is the number
//    of strong-neighbour sets it appears in; candidates are taken lowest
//    priority first, ties in order of their last (re)insertion, a
//    decrement moving the candidate behind the others of its new priority;
//  * pairing (find_pair, pairwise_agg, 208-270): the selected row pairs with
//    the remaining candidate of smallest a_rc in its row (first in CSR
//    order on ties) when that candidate is a strong neighbour, else stays
//    alone; the priorities of the strong neighbours of both are decremented;
//  * double (triple) pairwise (317-345): a second (third) pass on the
//    auxiliary matrix of the first pass's aggregates, without the dominance
//    check; the auxiliary matrix reads its aggregate ids through a
//    std::map whose operator[] gives id 0 to columns of rows outside every
//    aggregate (create_aux, 280-303), and that is kept.
// The priority queue is a bucket per priority with the candidates in
// insertion order -- the order of the reference's multimap, in O(1) per step.
#include "fsx_amg_setup.hpp"
#include "fsx_internal.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <vector>

namespace fsx {

// ---- device kernels -----------------------------------------------------------

// y[c] = sum over c's fine rows (ascending) of r[i], from 0 (restrict::ave)
template<class I>
__global__ __launch_bounds__(256) void amg_restrict_kernel(int64_t nc, const I * __restrict__ mptr,
                                                           const I * __restrict__ midx,
                                                           const double * __restrict__ r,
                                                           double * __restrict__ y) {
	const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (c >= nc) return;
	double s = 0;
	for (I q = mptr[c]; q < mptr[c + 1]; ++q) s += r[midx[q]];
	y[c] = s;
}

// x += P xc: x[i] + xc[agg[i]], rows outside every aggregate add the
// correction vector's untouched zero (prolong::inject, then x.add)
template<class I>
__global__ __launch_bounds__(256) void amg_prolong_add_kernel(int64_t n, const I * __restrict__ aggt,
                                                              const double * __restrict__ xc,
                                                              double * __restrict__ x) {
	const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (i >= n) return;
	const I a = aggt[i];
	x[i] = x[i] + (a >= 0 ? xc[a] : 0.0);
}

// z = a x + b y (linear_sum) and z = a x + y (axpy), elementwise
__global__ __launch_bounds__(256) void amg_lsum_kernel(int64_t n, double a, const double * __restrict__ x,
                                                       double b, const double * __restrict__ y,
                                                       double * __restrict__ z, int axpy) {
	const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (i >= n) return;
	z[i] = axpy ? a * x[i] + y[i] : a * x[i] + b * y[i];
}

} // namespace fsx

using namespace fsx;

struct fsx_amg {
	struct level {
		fsx_csr * A = nullptr;      // level 0: the caller's matrix (not owned)
		bool owned = false;
		fsx_gs * G = nullptr;       // smoother (all but the coarsest level)
		int64_t n = 0, nnz = 0;
		// map to the next coarser level: fine row -> aggregate (or -1), and
		// each aggregate's fine rows ascending
		std::unique_ptr<dbuf> aggt, mptr, midx;
		std::vector<int64_t> aggt_host;
		fsx_vec *res = nullptr, *rhs = nullptr, *sol = nullptr, *c = nullptr, *v = nullptr,
		        *bt = nullptr, *d = nullptr, *w = nullptr;
	};
	fsx_ctx * ctx = nullptr;
	fsx_amg_settings s{};
	std::vector<level> lv;
	dense_lu lu;
	std::vector<double> hb, hx; // coarse solve staging
	~fsx_amg() {
		for (auto & L : lv) {
			if (L.G) fsx_gs_destroy(L.G);
			if (L.owned && L.A) fsx_csr_destroy(L.A);
			for (fsx_vec * p : {L.res, L.rhs, L.sol, L.c, L.v, L.bt, L.d, L.w})
				if (p) fsx_vec_release(p);
		}
	}
};

namespace {

void rc(int code) {
	if (code) throw error(code, get_error());
}

hcsr download(fsx_ctx * ctx, const fsx_csr * A) {
	hcsr H;
	H.n = A->nrows;
	const size_t isz = A->it == FSX_I32 ? 4 : 8;
	std::vector<char> hrp((A->nrows + 1) * isz), hci(std::max<int64_t>(1, A->nnz) * isz);
	H.v.resize(A->nnz);
	FSX_HIP(hipMemcpyAsync(hrp.data(), A->rowptr(), hrp.size(), hipMemcpyDeviceToHost, ctx->stream));
	if (A->nnz) {
		FSX_HIP(hipMemcpyAsync(hci.data(), A->colind(), A->nnz * isz, hipMemcpyDeviceToHost,
		                       ctx->stream));
		FSX_HIP(hipMemcpyAsync(H.v.data(), A->vals->p, A->nnz * sizeof(double),
		                       hipMemcpyDeviceToHost, ctx->stream));
	}
	FSX_HIP(hipStreamSynchronize(ctx->stream));
	H.rp.resize(A->nrows + 1);
	H.ci.resize(A->nnz);
	for (int64_t i = 0; i <= A->nrows; ++i)
		H.rp[i] = isz == 4 ? ((int32_t *)hrp.data())[i] : ((int64_t *)hrp.data())[i];
	for (int64_t e = 0; e < A->nnz; ++e)
		H.ci[e] = isz == 4 ? ((int32_t *)hci.data())[e] : ((int64_t *)hci.data())[e];
	return H;
}

template<class I>
std::unique_ptr<dbuf> upload(fsx_ctx * ctx, const std::vector<int64_t> & h) {
	std::vector<I> t(h.begin(), h.end());
	auto d = std::make_unique<dbuf>(std::max<size_t>(1, t.size()) * sizeof(I));
	if (!t.empty())
		FSX_HIP(hipMemcpyAsync(d->p, t.data(), t.size() * sizeof(I), hipMemcpyHostToDevice,
		                       ctx->stream));
	FSX_HIP(hipStreamSynchronize(ctx->stream));
	return d;
}

fsx_vec * newvec(fsx_ctx * ctx, int64_t n) {
	fsx_vec * v = nullptr;
	rc(fsx_vec_create(ctx, (size_t)n, FSX_F64, &v));
	launch_set(ctx, 0.0, v);
	return v;
}

unsigned blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

// one level's smoothing: nrelax symmetric hybrid Gauss-Seidel sweeps, or
// nrelax weighted Jacobi sweeps
void smooth(fsx_amg & M, size_t l, const fsx_vec * b, fsx_vec * x) {
	auto & L = M.lv[l];
	if (M.s.smoother == 0) {
		rc(fsx_gs_relax(M.ctx, L.G, 2, M.s.nrelax, b, x));
		return;
	}
	rc(fsx_jacobi_relax(M.ctx, L.A, M.s.jacobi_weight, M.s.nrelax, b, x));
}

void restrict_to(fsx_amg & M, size_t l, const fsx_vec * r, fsx_vec * y) {
	const auto & F = M.lv[l];
	const int64_t nc = M.lv[l + 1].n;
	if (F.A->it == FSX_I32)
		hipLaunchKernelGGL(amg_restrict_kernel<int32_t>, dim3(blocks(nc)), dim3(256), 0, M.ctx->stream,
		                   nc, (const int32_t *)F.mptr->p, (const int32_t *)F.midx->p,
		                   (const double *)r->ptr(), (double *)y->ptr());
	else
		hipLaunchKernelGGL(amg_restrict_kernel<int64_t>, dim3(blocks(nc)), dim3(256), 0, M.ctx->stream,
		                   nc, (const int64_t *)F.mptr->p, (const int64_t *)F.midx->p,
		                   (const double *)r->ptr(), (double *)y->ptr());
	FSX_HIP(hipGetLastError());
}

void prolong_add(fsx_amg & M, size_t l, const fsx_vec * xc, fsx_vec * x) {
	const auto & F = M.lv[l];
	if (F.A->it == FSX_I32)
		hipLaunchKernelGGL(amg_prolong_add_kernel<int32_t>, dim3(blocks(F.n)), dim3(256), 0,
		                   M.ctx->stream, F.n, (const int32_t *)F.aggt->p, (const double *)xc->ptr(),
		                   (double *)x->ptr());
	else
		hipLaunchKernelGGL(amg_prolong_add_kernel<int64_t>, dim3(blocks(F.n)), dim3(256), 0,
		                   M.ctx->stream, F.n, (const int64_t *)F.aggt->p, (const double *)xc->ptr(),
		                   (double *)x->ptr());
	FSX_HIP(hipGetLastError());
}

void lsum(fsx_amg & M, int64_t n, double a, const fsx_vec * x, double b, const fsx_vec * y,
          fsx_vec * z, bool axpy) {
	hipLaunchKernelGGL(amg_lsum_kernel, dim3(blocks(n)), dim3(256), 0, M.ctx->stream, n, a,
	                   (const double *)x->ptr(), b, (const double *)y->ptr(), (double *)z->ptr(),
	                   axpy ? 1 : 0);
	FSX_HIP(hipGetLastError());
}

void coarse_solve(fsx_amg & M, const fsx_vec * b, fsx_vec * x) {
	const int64_t n = M.lu.n;
	M.hb.resize((size_t)n);
	M.hx.resize((size_t)n);
	rc(fsx_vec_download(M.ctx, b, M.hb.data(), (size_t)n));
	M.lu.solve(M.hb.data(), M.hx.data());
	rc(fsx_vec_upload(M.ctx, x, M.hx.data(), (size_t)n));
}

// cycle kinds: 0 v (ncycle, n = 1), 1 relaxed W (tau 1.75), 2 relaxed
// kappa (tau 1.75), 3 kappa-K (cycle.hh:43-272)
void cycle(fsx_amg & M, size_t l, const fsx_vec * b, fsx_vec * x, int kappa) {
	auto & L = M.lv[l];
	auto & C = M.lv[l + 1];
	const size_t coarse_lvl = M.lv.size() - 1;
	const int kind = M.s.cycle;
	const double tau = 1.75;
	smooth(M, l, b, x);
	rc(fsx_csr_residual(M.ctx, L.A, b, x, L.res));
	restrict_to(M, l, L.res, C.rhs);
	launch_set(M.ctx, 0.0, C.sol);
	if (l + 1 == coarse_lvl)
		coarse_solve(M, C.rhs, C.sol);
	else if (kind == 0)
		cycle(M, l + 1, C.rhs, C.sol, kappa);
	else if (kind == 1 || (kind == 2 && kappa > 1)) {
		const int kn = kind == 2 ? kappa - 1 : kappa;
		launch_set(M.ctx, 0.0, C.c);
		cycle(M, l + 1, C.rhs, C.c, kappa);
		rc(fsx_csr_spmv(M.ctx, C.A, C.c, C.v));
		lsum(M, C.n, -tau, C.v, 0.0, C.rhs, C.bt, true);
		launch_set(M.ctx, 0.0, C.d);
		cycle(M, l + 1, C.bt, C.d, kn);
		lsum(M, C.n, tau, C.c, tau, C.d, C.sol, false);
	}
	else if (kind == 2)
		cycle(M, l + 1, C.rhs, C.sol, kappa);
	else if (kappa > 1) { // kappa-K
		launch_set(M.ctx, 0.0, C.c);
		cycle(M, l + 1, C.rhs, C.c, kappa);
		rc(fsx_csr_spmv(M.ctx, C.A, C.c, C.v));
		const double rho1 = reduce_host(M.ctx, RED_DOT, C.c, C.v);
		const double alpha1 = reduce_host(M.ctx, RED_DOT, C.c, C.rhs);
		const double tau1 = alpha1 / rho1;
		lsum(M, C.n, -tau1, C.v, 0.0, C.rhs, C.bt, true);
		const double nbt = std::sqrt(reduce_host(M.ctx, RED_L2SQ, C.bt, nullptr));
		const double nb = std::sqrt(reduce_host(M.ctx, RED_L2SQ, C.rhs, nullptr));
		if (nbt < M.s.ktol * nb)
			lsum(M, C.n, tau1, C.c, 0.0, C.sol, C.sol, true);
		else {
			launch_set(M.ctx, 0.0, C.d);
			cycle(M, l + 1, C.bt, C.d, kappa - 1);
			rc(fsx_csr_spmv(M.ctx, C.A, C.d, C.w));
			const double gamma = reduce_host(M.ctx, RED_DOT, C.d, C.v);
			const double beta = reduce_host(M.ctx, RED_DOT, C.d, C.w);
			const double alpha2 = reduce_host(M.ctx, RED_DOT, C.d, C.bt);
			const double rho2 = beta - ((gamma * gamma) / rho1);
			const double tau2 = tau1 - (gamma * alpha2) / (rho1 * rho2);
			const double tau3 = alpha2 / rho2;
			lsum(M, C.n, tau2, C.c, tau3, C.d, C.sol, false);
		}
	}
	else {
		launch_set(M.ctx, 0.0, C.sol);
		cycle(M, l + 1, C.rhs, C.sol, kappa);
	}
	prolong_add(M, l, C.sol, x);
	smooth(M, l, b, x);
}

} // namespace

extern "C" {

int fsx_amg_create(fsx_ctx * ctx, const fsx_csr * A, const fsx_amg_settings * s,
                   fsx_amg ** out) {
	return guard([&] {
		FSX_REQUIRE(ctx && A && s && out, "null argument");
		FSX_REQUIRE(A->nrows == A->ncols && A->nrows > 0, "amg: square, non-empty matrix");
		FSX_REQUIRE(A->vt == FSX_F64, "amg: FP64 matrix");
		FSX_REQUIRE(s->max_levels >= 1, "amg: max_levels >= 1");
		FSX_REQUIRE(s->pairwise_passes == 2 || s->pairwise_passes == 3,
		            "amg: pairwise passes 2 or 3");
		FSX_REQUIRE(s->cycle >= 0 && s->cycle <= 3, "amg: cycle 0 v, 1 relaxed-w, "
		                                             "2 relaxed-kappa, 3 kappa-k");
		FSX_REQUIRE(s->nrelax >= 0 && s->maxiter >= 0, "amg: nrelax, maxiter >= 0");
		FSX_REQUIRE(s->smoother == 0 || s->smoother == 1,
		            "amg: smoother 0 (hybrid Gauss-Seidel) or 1 (weighted Jacobi)");
		auto M = std::make_unique<fsx_amg>();
		M->ctx = ctx;
		M->s = *s;
		hcsr H = download(ctx, A);
		for (int64_t e = 0; e < (int64_t)H.ci.size(); ++e)
			FSX_REQUIRE(H.ci[e] >= 0 && H.ci[e] < H.n, "amg: column index out of range");
		M->lv.emplace_back();
		M->lv[0].A = const_cast<fsx_csr *>(A);
		M->lv[0].n = H.n;
		M->lv[0].nnz = (int64_t)H.ci.size();
		// bound_solver::setup (ua.hh:110-139): coarsen until a level has
		// fewer than min_coarse rows or max_levels levels exist
		const bool i32 = A->it == FSX_I32;
		H = coarsen_levels(std::move(H), s->beta, s->pairwise_passes, s->max_levels,
		                   s->min_coarse,
		                   [&](int, const hcsr &, const aggregates & agg,
		                       const std::vector<int64_t> & aggt, const hcsr & Cm) {
			auto & F = M->lv.back();
			F.aggt_host = aggt;
			// each aggregate's fine rows, ascending (restrict::ave sums them
			// in fine-row order)
			std::vector<int64_t> mptr(agg.size() + 1, 0), midx;
			midx.reserve(agg.idx.size());
			for (size_t a = 0; a < agg.size(); ++a) {
				const size_t b0 = midx.size();
				midx.insert(midx.end(), agg[a].begin(), agg[a].end());
				std::sort(midx.begin() + (std::ptrdiff_t)b0, midx.end());
				mptr[a + 1] = (int64_t)midx.size();
			}
			F.aggt = i32 ? upload<int32_t>(ctx, F.aggt_host) : upload<int64_t>(ctx, F.aggt_host);
			F.mptr = i32 ? upload<int32_t>(ctx, mptr) : upload<int64_t>(ctx, mptr);
			F.midx = i32 ? upload<int32_t>(ctx, midx) : upload<int64_t>(ctx, midx);
			fsx_amg::level C;
			C.n = Cm.n;
			C.nnz = (int64_t)Cm.ci.size();
			C.owned = true;
			if (i32) {
				std::vector<int32_t> rp(Cm.rp.begin(), Cm.rp.end()), ci(Cm.ci.begin(), Cm.ci.end());
				rc(fsx_csr_create(ctx, Cm.n, Cm.n, C.nnz, FSX_I32, rp.data(), ci.data(), FSX_F64,
				                  Cm.v.data(), &C.A));
			}
			else
				rc(fsx_csr_create(ctx, Cm.n, Cm.n, C.nnz, FSX_I64, Cm.rp.data(), Cm.ci.data(),
				                  FSX_F64, Cm.v.data(), &C.A));
			// row sums in CSR order on every level (bit-identical residuals)
			int kind = 0, lanes = 0;
			rc(fsx_csr_plan(C.A, &kind, &lanes));
			if (kind == 1) rc(fsx_csr_set_plan(C.A, 2, 0));
			M->lv.push_back(std::move(C));
		});
		for (size_t l = 0; l < M->lv.size(); ++l) {
			auto & L = M->lv[l];
			if (l + 1 < M->lv.size() && s->smoother == 0) rc(fsx_gs_create(ctx, L.A, 1, &L.G));
			L.res = newvec(ctx, L.n);
			if (l > 0) {
				L.rhs = newvec(ctx, L.n);
				L.sol = newvec(ctx, L.n);
				L.c = newvec(ctx, L.n);
				L.v = newvec(ctx, L.n);
				L.bt = newvec(ctx, L.n);
				L.d = newvec(ctx, L.n);
				L.w = newvec(ctx, L.n);
			}
		}
		FSX_REQUIRE(H.n <= 4096, "amg: coarsest level above 4096 rows (raise max_levels)");
		M->lu.factor(H);
		ctx_retain(ctx);
		*out = M.release();
	});
}

int fsx_amg_levels(const fsx_amg * M, int * nlevels, int64_t * rows, int64_t * nnz, int cap) {
	return guard([&] {
		FSX_REQUIRE(M && nlevels, "null argument");
		*nlevels = (int)M->lv.size();
		for (int l = 0; l < std::min<int>(cap, (int)M->lv.size()); ++l) {
			if (rows) rows[l] = M->lv[l].n;
			if (nnz) nnz[l] = M->lv[l].nnz;
		}
	});
}

int fsx_amg_aggregates(const fsx_amg * M, int level, int64_t * aggt) {
	return guard([&] {
		FSX_REQUIRE(M && aggt, "null argument");
		FSX_REQUIRE(level >= 0 && level + 1 < (int)M->lv.size(), "amg: no such fine level");
		std::copy(M->lv[level].aggt_host.begin(), M->lv[level].aggt_host.end(), aggt);
	});
}

int fsx_amg_level_matrix(const fsx_amg * M, int level, fsx_csr ** A) {
	return guard([&] {
		FSX_REQUIRE(M && A, "null argument");
		FSX_REQUIRE(level >= 0 && level < (int)M->lv.size(), "amg: no such level");
		*A = M->lv[level].A;
	});
}

int fsx_amg_apply(fsx_ctx * ctx, fsx_amg * M, const fsx_vec * b, fsx_vec * x, int * cycles,
                  double * res) {
	return guard([&] {
		FSX_REQUIRE(ctx && M && M->ctx == ctx, "null argument or another context");
		check_vec(b, "b");
		check_vec(x, "x");
		FSX_REQUIRE((int64_t)b->n == M->lv[0].n && (int64_t)x->n == M->lv[0].n,
		            "amg: vector length");
		FSX_REQUIRE(b->dt == FSX_F64 && x->dt == FSX_F64, "amg: FP64 vectors");
		FSX_REQUIRE(b->ptr() != x->ptr(), "amg: b and x must differ");
		auto & L0 = M->lv[0];
		// bound_solver::apply (ua.hh:141-172).  The residual norms only decide
		// whether another cycle runs and what `res` reports: the one before
		// the first cycle is needed only when no cycle runs (it is otherwise
		// overwritten), the one after a cycle only when another may follow or
		// the caller asks for `res`; the others are skipped (x is the same).
		launch_set(ctx, 0.0, x);
		const int maxit = M->s.maxiter;
		double rnorm = 0.0;
		if (maxit <= 0) {
			rc(fsx_csr_residual(ctx, L0.A, b, x, L0.res));
			rnorm = std::sqrt(reduce_host(ctx, RED_L2SQ, L0.res, nullptr));
		}
		int done = 0;
		for (int i = 0; i < maxit; ++i) {
			if (M->lv.size() == 1) coarse_solve(*M, b, x);
			else cycle(*M, 0, b, x, M->s.kappa);
			++done;
			if (i + 1 == maxit && !res) break;
			rc(fsx_csr_residual(ctx, L0.A, b, x, L0.res));
			rnorm = std::sqrt(reduce_host(ctx, RED_L2SQ, L0.res, nullptr));
			if (rnorm <= (double)M->s.atol) break;
		}
		if (cycles) *cycles = done;
		if (res) *res = rnorm;
	});
}

int fsx_amg_destroy(fsx_amg * M) {
	return guard([&] {
		if (!M) return;
		fsx_ctx * c = M->ctx;
		(void)hipStreamSynchronize(c->stream);
		delete M;
		ctx_release(c);
	});
}

} // extern "C"
