// fsx_smooth.hip -- hybrid Gauss-Seidel relaxation on the device.
//
// Reference: op::hybrid_gs (flecsolve/solvers/mg/gs.hh:40-130).  One sweep
// runs on every colour (a contiguous block of rows owned by one rank) at
// once; inside a colour the rows are visited one after another, in order
// (forward) or reverse order (backward), and row r becomes
//     diag = a_rr,  rsum = sum over the colour's own columns c != r, in CSR
//     order, of a_rc x_c, then + sum over the other colours' columns, in CSR
//     order, of a_rc x_c;   x_r = (1. / diag) (b_r - rsum)
// where x_c of the own colour is already this sweep's value when c was
// visited before r, and every other x_c is the value before the sweep (the
// ghost copy).  symmetric = forward then backward; `nrelax` repetitions.
// Colours follow FleCSI's equal_map(n, P): colour i owns rows
// [i q + min(i, m), (i+1) q + min(i+1, m)), q = n / P, m = n % P.
//
// On the GPU a row needs only the rows of its own colour that are visited
// before it and that it couples to.  The host groups rows into levels
// (level(r) = 1 + the largest level of those rows, 0 if none); each level is
// one launch with a thread per row, rows of a level in ascending order.  A
// thread reads x for its own colour's earlier rows (finished in lower
// levels) and a snapshot of x taken at the start of the sweep for everything
// else, and does the reference's arithmetic in the reference's order, so the
// result is bit-identical to the sequential sweep.  The number of levels is
// the length of the longest dependency chain (3n - 2 for the natural-order
// 7-point stencil on n^3): a sweep is a chain of that many short launches,
// captured once per (b, x, sweep kind, nrelax) in a HIP graph.
#include "fsx_internal.hpp"

#include <algorithm>
#include <map>
#include <vector>

namespace fsx {

// FleCSI's equal_map(n, P): block i starts at i q + min(i, m).
struct colour_map {
	int64_t q = 1, m = 0, big = 0; // big = m (q + 1): rows of the larger blocks
	__host__ __device__ int64_t of(int64_t j) const {
		return j < big ? j / (q + 1) : m + (j - big) / q;
	}
};

// One pass over row r's entries in CSR order, restricted to own-colour
// columns (OWN) or to the other colours' columns (!OWN; COLOURS only).
// Entries go in tiles of kGsTile: every load of a tile is issued before the
// first use, then the tile is consumed in order (a loop that loaded one
// entry, gathered its x and added it before the next paid the memory
// latency once per entry).
constexpr int kGsTile = 8;

template<class I, bool FWD, bool COLOURS, bool OWN>
__device__ __forceinline__ void gs_row_pass(int64_t r, I lo, I hi, int64_t cr,
                                            const I * __restrict__ ci,
                                            const double * __restrict__ va, const double * x,
                                            const double * __restrict__ xold,
                                            const colour_map & cm, double & diag,
                                            double & rsum) {
	for (I e0 = lo; e0 < hi; e0 += kGsTile) {
		int64_t c[kGsTile];
		double v[kGsTile], xv[kGsTile];
		bool pr[kGsTile];
#pragma unroll
		for (int k = 0; k < kGsTile; ++k) {
			const I e = min((I)(e0 + k), (I)(hi - 1));
			c[k] = ci[e];
			v[k] = va[e];
			pr[k] = e0 + k < hi;
			if (COLOURS) pr[k] = pr[k] && ((cm.of(c[k]) == cr) == OWN);
		}
#pragma unroll
		for (int k = 0; k < kGsTile; ++k) {
			// x only for present entries of rows finished in lower levels;
			// everything else (absent slots included) reads the snapshot
			const bool visited = OWN && pr[k] && (FWD ? c[k] < r : c[k] > r);
			xv[k] = (visited ? x : xold)[c[k]];
		}
#pragma unroll
		for (int k = 0; k < kGsTile; ++k) {
			if (!pr[k]) continue;
			if (OWN && c[k] == r) diag = v[k];
			else rsum += v[k] * xv[k];
		}
	}
}

template<class I, bool FWD, bool COLOURS>
__global__ __launch_bounds__(256) void gs_level_kernel(
	const I * __restrict__ rp, const I * __restrict__ ci, const double * __restrict__ va,
	const I * __restrict__ rows, int64_t count, const double * __restrict__ b, double * x,
	const double * __restrict__ xold, colour_map cm) {
	const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (t >= count) return;
	const int64_t r = rows[t];
	const I lo = rp[r], hi = rp[r + 1];
	const double br = b[r];
	const int64_t cr = COLOURS ? cm.of(r) : 0;
	double diag = 0.;
	double rsum = 0.;
	// the colour's own block (gs.hh:94-104), CSR order
	gs_row_pass<I, FWD, COLOURS, true>(r, lo, hi, cr, ci, va, x, xold, cm, diag, rsum);
	// the other colours' columns (gs.hh:105-107), values before the sweep
	if (COLOURS)
		gs_row_pass<I, FWD, COLOURS, false>(r, lo, hi, cr, ci, va, x, xold, cm, diag, rsum);
	const double dinv = 1. / diag;
	x[r] = dinv * (br - rsum);
}

__global__ __launch_bounds__(256) void gs_snapshot_kernel(size_t n, const double * __restrict__ x,
                                                          double * __restrict__ xold) {
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) xold[i] = x[i];
}

} // namespace fsx

using namespace fsx;

struct fsx_gs {
	fsx_ctx * ctx = nullptr;
	const fsx_csr * A = nullptr;
	int ncolours = 1;
	colour_map cm;
	// rows sorted by (level, row) and level offsets, forward and backward
	std::unique_ptr<dbuf> rows_f, rows_b, xold;
	std::vector<int64_t> lev_f, lev_b;
	struct graph {
		hipGraph_t g = nullptr;
		hipGraphExec_t ge = nullptr;
	};
	std::map<std::vector<const void *>, graph> graphs;
	~fsx_gs() {
		for (auto & kv : graphs) {
			if (kv.second.ge) (void)hipGraphExecDestroy(kv.second.ge);
			if (kv.second.g) (void)hipGraphDestroy(kv.second.g);
		}
	}
};

namespace {

// Levels of one sweep direction and the rows grouped by level (counting
// sort: ascending row inside a level).
void schedule(const std::vector<int64_t> & rp, const std::vector<int64_t> & ci,
              const colour_map & cm, bool colours, bool fwd, std::vector<int64_t> & ptr,
              std::vector<int64_t> & order) {
	const int64_t n = (int64_t)rp.size() - 1;
	std::vector<int32_t> lev(n, 0);
	int32_t maxl = -1;
	for (int64_t s = 0; s < n; ++s) {
		const int64_t r = fwd ? s : n - 1 - s;
		const int64_t cr = colours ? cm.of(r) : 0;
		int32_t l = 0;
		for (int64_t e = rp[r]; e < rp[r + 1]; ++e) {
			const int64_t c = ci[e];
			if ((fwd ? c < r : c > r) && (!colours || cm.of(c) == cr))
				l = std::max(l, lev[c] + 1);
		}
		lev[r] = l;
		maxl = std::max(maxl, l);
	}
	ptr.assign((size_t)maxl + 2, 0);
	for (int64_t r = 0; r < n; ++r) ptr[lev[r] + 1]++;
	for (int32_t l = 0; l <= maxl; ++l) ptr[l + 1] += ptr[l];
	std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
	order.resize(n);
	for (int64_t r = 0; r < n; ++r) order[fill[lev[r]]++] = r;
}

template<class I>
std::unique_ptr<dbuf> upload_rows(fsx_ctx * ctx, const std::vector<int64_t> & order) {
	std::vector<I> h(order.begin(), order.end());
	auto d = std::make_unique<dbuf>(std::max<size_t>(1, h.size()) * sizeof(I));
	if (!h.empty())
		FSX_HIP(hipMemcpyAsync(d->p, h.data(), h.size() * sizeof(I), hipMemcpyHostToDevice,
		                       ctx->stream));
	FSX_HIP(hipStreamSynchronize(ctx->stream)); // h goes out of scope
	return d;
}

template<class I>
void enqueue_sweep(const fsx_gs & G, bool fwd, const double * b, double * x) {
	const fsx_csr * A = G.A;
	hipStream_t s = G.ctx->stream;
	const size_t n = (size_t)A->nrows;
	hipLaunchKernelGGL(gs_snapshot_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
	                   n, (const double *)x, (double *)G.xold->p);
	const std::vector<int64_t> & ptr = fwd ? G.lev_f : G.lev_b;
	const I * rows = (const I *)(fwd ? G.rows_f : G.rows_b)->p;
	const I * rp = (const I *)A->rowptr();
	const I * ci = (const I *)A->colind();
	const double * va = (const double *)A->vals->p;
	const double * xo = (const double *)G.xold->p;
	for (size_t l = 0; l + 1 < ptr.size(); ++l) {
		const int64_t cnt = ptr[l + 1] - ptr[l];
		const unsigned g = (unsigned)((cnt + 255) / 256);
		const I * rl = rows + ptr[l];
		if (G.ncolours > 1) {
			if (fwd)
				hipLaunchKernelGGL((gs_level_kernel<I, true, true>), dim3(g), dim3(256), 0, s, rp,
				                   ci, va, rl, cnt, b, x, xo, G.cm);
			else
				hipLaunchKernelGGL((gs_level_kernel<I, false, true>), dim3(g), dim3(256), 0, s, rp,
				                   ci, va, rl, cnt, b, x, xo, G.cm);
		}
		else {
			if (fwd)
				hipLaunchKernelGGL((gs_level_kernel<I, true, false>), dim3(g), dim3(256), 0, s,
				                   rp, ci, va, rl, cnt, b, x, xo, G.cm);
			else
				hipLaunchKernelGGL((gs_level_kernel<I, false, false>), dim3(g), dim3(256), 0, s,
				                   rp, ci, va, rl, cnt, b, x, xo, G.cm);
		}
	}
	FSX_HIP(hipGetLastError());
}

// sweep kinds: 0 forward, 1 backward, 2 symmetric (mg::relax_sweep)
void enqueue_relax(const fsx_gs & G, int sweep, int nrelax, const double * b, double * x) {
	for (int it = 0; it < nrelax; ++it) {
		auto run = [&](bool fwd) {
			if (G.A->it == FSX_I32) enqueue_sweep<int32_t>(G, fwd, b, x);
			else enqueue_sweep<int64_t>(G, fwd, b, x);
		};
		if (sweep == 0) run(true);
		else if (sweep == 1) run(false);
		else {
			run(true);
			run(false);
		}
	}
}

bool gs_graphs() {
	static bool g = [] {
		const char * e = getenv("FSX_GS_GRAPH");
		return !(e && e[0] == '0');
	}();
	return g;
}

} // namespace

extern "C" {

int fsx_gs_create(fsx_ctx * ctx, const fsx_csr * A, int ncolours, fsx_gs ** out) {
	return guard([&] {
		FSX_REQUIRE(ctx && A && out, "null argument");
		FSX_REQUIRE(A->nrows == A->ncols, "hybrid_gs needs a square matrix");
		FSX_REQUIRE(A->vt == FSX_F64, "hybrid_gs: FP64 matrix");
		FSX_REQUIRE(ncolours >= 1 && (A->nrows == 0 || ncolours <= A->nrows),
		            "hybrid_gs: 1 <= colours <= rows");
		const int64_t n = A->nrows;
		const size_t isz = A->it == FSX_I32 ? 4 : 8;
		std::vector<int64_t> rp(n + 1), ci(A->nnz);
		{
			std::vector<char> hrp((n + 1) * isz), hci(std::max<int64_t>(1, A->nnz) * isz);
			FSX_HIP(hipMemcpyAsync(hrp.data(), A->rowptr(), hrp.size(), hipMemcpyDeviceToHost,
			                       ctx->stream));
			if (A->nnz)
				FSX_HIP(hipMemcpyAsync(hci.data(), A->colind(), A->nnz * isz,
				                       hipMemcpyDeviceToHost, ctx->stream));
			FSX_HIP(hipStreamSynchronize(ctx->stream));
			for (int64_t i = 0; i <= n; ++i)
				rp[i] = isz == 4 ? ((int32_t *)hrp.data())[i] : ((int64_t *)hrp.data())[i];
			for (int64_t e = 0; e < A->nnz; ++e)
				ci[e] = isz == 4 ? ((int32_t *)hci.data())[e] : ((int64_t *)hci.data())[e];
		}
		for (int64_t e = 0; e < A->nnz; ++e)
			FSX_REQUIRE(ci[e] >= 0 && ci[e] < n, "hybrid_gs: column index out of range");
		auto G = std::make_unique<fsx_gs>();
		G->ctx = ctx;
		G->A = A;
		G->ncolours = ncolours;
		if (n > 0) {
			G->cm.q = n / ncolours;
			G->cm.m = n % ncolours;
			G->cm.big = G->cm.m * (G->cm.q + 1);
		}
		std::vector<int64_t> order;
		schedule(rp, ci, G->cm, ncolours > 1, true, G->lev_f, order);
		G->rows_f = isz == 4 ? upload_rows<int32_t>(ctx, order) : upload_rows<int64_t>(ctx, order);
		schedule(rp, ci, G->cm, ncolours > 1, false, G->lev_b, order);
		G->rows_b = isz == 4 ? upload_rows<int32_t>(ctx, order) : upload_rows<int64_t>(ctx, order);
		G->xold = std::make_unique<dbuf>(std::max<int64_t>(1, n) * sizeof(double));
		ctx_retain(ctx);
		*out = G.release();
	});
}

int fsx_gs_levels(const fsx_gs * G, int64_t * forward, int64_t * backward) {
	return guard([&] {
		FSX_REQUIRE(G, "null smoother");
		if (forward) *forward = (int64_t)G->lev_f.size() - 1;
		if (backward) *backward = (int64_t)G->lev_b.size() - 1;
	});
}

int fsx_gs_relax(fsx_ctx * ctx, fsx_gs * G, int sweep, int nrelax, const fsx_vec * b,
                 fsx_vec * x) {
	return guard([&] {
		FSX_REQUIRE(ctx && G && G->ctx == ctx, "null argument or another context");
		FSX_REQUIRE(sweep >= 0 && sweep <= 2, "sweep: 0 forward, 1 backward, 2 symmetric");
		FSX_REQUIRE(nrelax >= 0, "nrelax >= 0");
		check_vec(b, "b");
		check_vec(x, "x");
		FSX_REQUIRE((int64_t)b->n == G->A->nrows && (int64_t)x->n == G->A->nrows,
		            "hybrid_gs: vector length");
		FSX_REQUIRE(b->dt == FSX_F64 && x->dt == FSX_F64, "hybrid_gs: FP64 vectors");
		FSX_REQUIRE(b->ptr() != x->ptr() || b->n == 0, "hybrid_gs: b and x must differ");
		if (nrelax == 0 || G->A->nrows == 0) return;
		const double * bp = (const double *)b->ptr();
		double * xp = (double *)x->ptr();
		if (!gs_graphs()) {
			enqueue_relax(*G, sweep, nrelax, bp, xp);
			return;
		}
		std::vector<const void *> key{bp, xp, (const void *)(intptr_t)sweep,
		                              (const void *)(intptr_t)nrelax, G->A->vals->p};
		auto it = G->graphs.find(key);
		if (it == G->graphs.end()) {
			if (G->graphs.size() >= 16) {
				for (auto & kv : G->graphs) {
					(void)hipGraphExecDestroy(kv.second.ge);
					(void)hipGraphDestroy(kv.second.g);
				}
				G->graphs.clear();
			}
			fsx_gs::graph gr;
			FSX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
			try {
				enqueue_relax(*G, sweep, nrelax, bp, xp);
			}
			catch (...) {
				hipGraph_t junk = nullptr;
				(void)hipStreamEndCapture(ctx->stream, &junk);
				if (junk) (void)hipGraphDestroy(junk);
				throw;
			}
			FSX_HIP(hipStreamEndCapture(ctx->stream, &gr.g));
			FSX_HIP(hipGraphInstantiate(&gr.ge, gr.g, nullptr, nullptr, 0));
			it = G->graphs.emplace(key, gr).first;
		}
		FSX_HIP(hipGraphLaunch(it->second.ge, ctx->stream));
		ctx->launches += 1;
	});
}

int fsx_gs_destroy(fsx_gs * G) {
	return guard([&] {
		if (!G) return;
		fsx_ctx * c = G->ctx;
		(void)hipStreamSynchronize(c->stream);
		delete G;
		ctx_release(c);
	});
}

} // extern "C"
