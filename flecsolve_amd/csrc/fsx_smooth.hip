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
//
// Each level costs its chain of dependent loads, so each sweep direction
// keeps a level-ordered copy of what its rows read: per position (a row of
// a level, levels one after another) the row id, the diagonal value and the
// entry count, and per level an ELL block (entry j of the level's t-th row
// at base + j * count + t) of the off-diagonal entries in the order the
// reference adds them, each column tagged with where its x comes from
// (this sweep's x, or the snapshot: stored as -c - 1).  A thread loads all
// of that at once, then b and the x gathers, then stores: two dependent
// steps instead of four through row ids and row pointers.  The copy's
// values are refreshed from the matrix when its values change in place.
#include "fsx_host_threads.hpp"
#include "fsx_internal.hpp"

#include <algorithm>
#include <limits>
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

constexpr int kGsTile = 8;

// A row of a level in the level-ordered copy: its id, entry count, diagonal
// and its first kRowW entries (tagged columns and values).  A level's ELL
// block is `width` entries deep (padding: column tag -1, value 0), so all
// kRowW loads go out at once, before the entry count arrives.
constexpr int kRowW = 16;

template<class I>
struct gs_row {
	int64_t r;
	int len;
	double diag;
	I c[kRowW];
	double v[kRowW];
};

// Load row t of the level that starts at position `start` (cnt rows, ELL
// block at `base`, `width` entries deep).
template<class I>
__device__ __forceinline__ void gs_load_row(gs_row<I> & m, const I * __restrict__ srow,
                                            const int32_t * __restrict__ slen,
                                            const double * __restrict__ sdiag,
                                            const I * __restrict__ scol,
                                            const double * __restrict__ sval, int64_t start,
                                            int64_t t, int64_t cnt, int64_t base, int width) {
	const int64_t p = start + t;
	m.r = srow[p];
	m.len = slen[p];
	m.diag = sdiag[p];
	const I * cp = scol + base + t;
	const double * vp = sval + base + t;
#pragma unroll
	for (int k = 0; k < kRowW; ++k)
		if (k < width) {
			m.c[k] = cp[k * cnt];
			m.v[k] = vp[k * cnt];
		}
}

// Relax the row: b and the x gathers, the sum in entry order (entries past
// kRowW loaded in tiles of kGsTile), the update.
template<class I>
__device__ __forceinline__ void gs_finish_row(const gs_row<I> & m, const I * __restrict__ scol,
                                              const double * __restrict__ sval, int64_t t,
                                              int64_t cnt, int64_t base,
                                              const double * __restrict__ b, double * x,
                                              const double * __restrict__ xold) {
	const double br = b[m.r];
	double xv[kRowW];
#pragma unroll
	for (int k = 0; k < kRowW; ++k)
		if (k < m.len)
			xv[k] = (m.c[k] >= 0 ? x : xold)[m.c[k] >= 0 ? (int64_t)m.c[k] : -(int64_t)m.c[k] - 1];
	double rsum = 0.;
#pragma unroll
	for (int k = 0; k < kRowW; ++k)
		if (k < m.len) rsum += m.v[k] * xv[k];
	const I * cp = scol + base + t;
	const double * vp = sval + base + t;
	for (int j0 = kRowW; j0 < m.len; j0 += kGsTile) {
		I c[kGsTile];
		double v[kGsTile], xt[kGsTile];
#pragma unroll
		for (int k = 0; k < kGsTile; ++k) {
			const int64_t j = min(j0 + k, m.len - 1);
			c[k] = cp[j * cnt];
			v[k] = vp[j * cnt];
		}
#pragma unroll
		for (int k = 0; k < kGsTile; ++k)
			xt[k] = (c[k] >= 0 ? x : xold)[c[k] >= 0 ? (int64_t)c[k] : -(int64_t)c[k] - 1];
#pragma unroll
		for (int k = 0; k < kGsTile; ++k)
			if (j0 + k < m.len) rsum += v[k] * xt[k];
	}
	const double dinv = 1. / m.diag;
	x[m.r] = dinv * (br - rsum);
}

// One level: thread t relaxes the level's t-th row.
template<class I>
__global__ __launch_bounds__(256) void gs_level_kernel(
	const I * __restrict__ srow, const int32_t * __restrict__ slen,
	const double * __restrict__ sdiag, const I * __restrict__ scol,
	const double * __restrict__ sval, int64_t start, int64_t cnt, int64_t base, int width,
	const double * __restrict__ b, double * x, const double * __restrict__ xold) {
	const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (t >= cnt) return;
	gs_row<I> m;
	gs_load_row<I>(m, srow, slen, sdiag, scol, sval, start, t, cnt, base, width);
	gs_finish_row<I>(m, scol, sval, t, cnt, base, b, x, xold);
}

// A run of consecutive small levels (at most kRunThreads rows each) in one
// workgroup: the levels one after another, a barrier between (it makes this
// level's x stores visible to the whole workgroup before the next level
// reads them), thread t relaxing each level's t-th row.  The next level's
// row (id, count, diagonal, entries: nothing that depends on x) is loaded
// while this level's x gathers are in flight.  Measured on the coarse
// levels of a 256^3 aggregation hierarchy: ~1.7 us per level in a run
// against ~2.4 us per launch of a small level in a captured graph; levels
// of more than a few hundred rows keep a launch each (one workgroup would
// relax them row after row).  Level offsets, ELL bases and widths of the
// run are staged in shared memory first.
constexpr int kRunThreads = 512;
constexpr int kRunMaxLevels = 1024;
constexpr int64_t kRunRowsDefault = 256;

template<class I>
__global__ __launch_bounds__(kRunThreads) void gs_run_kernel(
	const I * __restrict__ srow, const int32_t * __restrict__ slen,
	const double * __restrict__ sdiag, const I * __restrict__ scol,
	const double * __restrict__ sval, const int64_t * __restrict__ lev,
	const int64_t * __restrict__ lbase, const int32_t * __restrict__ lwidth, int64_t l0,
	int nlev, const double * __restrict__ b, double * x, const double * __restrict__ xold) {
	__shared__ int64_t s_lev[kRunMaxLevels + 1], s_base[kRunMaxLevels];
	__shared__ int32_t s_w[kRunMaxLevels];
	for (int i = threadIdx.x; i <= nlev; i += kRunThreads) {
		s_lev[i] = lev[l0 + i];
		if (i < nlev) {
			s_base[i] = lbase[l0 + i];
			s_w[i] = lwidth[l0 + i];
		}
	}
	__syncthreads();
	const int64_t t = threadIdx.x;
	gs_row<I> cur, nxt;
	if (t < s_lev[1] - s_lev[0])
		gs_load_row<I>(cur, srow, slen, sdiag, scol, sval, s_lev[0], t, s_lev[1] - s_lev[0],
		               s_base[0], s_w[0]);
	for (int l = 0; l < nlev; ++l) {
		const int64_t cnt = s_lev[l + 1] - s_lev[l];
		if (l + 1 < nlev) {
			const int64_t c1 = s_lev[l + 2] - s_lev[l + 1];
			if (t < c1)
				gs_load_row<I>(nxt, srow, slen, sdiag, scol, sval, s_lev[l + 1], t, c1,
				               s_base[l + 1], s_w[l + 1]);
		}
		if (t < cnt) gs_finish_row<I>(cur, scol, sval, t, cnt, s_base[l], b, x, xold);
		__syncthreads();
		if (l + 1 < nlev) cur = nxt;
	}
}

// Refresh the copy's values from the matrix (entry index per slot, -1 for
// padding / a missing diagonal).
__global__ __launch_bounds__(256) void gs_gather_kernel(size_t n, const int64_t * __restrict__ from,
                                                        const double * __restrict__ vals,
                                                        double * __restrict__ out) {
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) {
		const int64_t e = from[i];
		out[i] = e >= 0 ? vals[e] : 0.0;
	}
}

__global__ __launch_bounds__(256) void gs_snapshot_kernel(size_t n, const double * __restrict__ x,
                                                          double * __restrict__ xold) {
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) xold[i] = x[i];
}

// ---- stencil matrices on a grid: tiles relaxed in shared memory ------------
//
// A one-colour matrix whose every entry couples a row to itself or to a face
// neighbour of its cell on an nx x ny x nz grid in natural order (offsets
// +-1, +-nx, +-nx ny; columns ascending in each row) is swept in tiles of
// kGridTile^3 cells.  A tile needs its faces' neighbours from the tiles
// before it only, so tiles on one diagonal I + J + K run at once (face
// neighbours always lie on the diagonals before and after: the halo holds
// finished values from the one and untouched values from the other), one
// launch per diagonal.  In a tile, one thread per cell; x for the tile and
// a one-cell halo sits in shared memory and the cells are relaxed diagonal
// by diagonal (i + j + k), a barrier between, each with the reference's
// arithmetic in CSR order: the sequential sweep's result, bit for bit, in
// nx/8 + ny/8 + nz/8 - 2 launches instead of nx + ny + nz - 2.
constexpr int kGridTile = 8;
constexpr int kGridHalo = kGridTile + 2;
constexpr int kGridLds = kGridHalo * kGridHalo * kGridHalo;
constexpr int kGridThreads = kGridTile * kGridTile * kGridTile;

struct grid_desc {
	int64_t nx, ny, nz, nxy; // grid; nxy = nx ny
	int ti, tj, tk;          // tiles per dimension
};

// Entry slots in column order: 0 -nxy, 1 -nx, 2 -1, 3 the diagonal, 4 +1,
// 5 +nx, 6 +nxy; gval[slot * n + r], bit slot of gmask[r] set when present.
__host__ __device__ inline int grid_slot(int64_t o, const grid_desc & g) {
	return o == -g.nxy ? 0 : o == -g.nx ? 1 : o == -1 ? 2 : o == 0 ? 3 : o == 1 ? 4
	     : o == g.nx ? 5 : 6;
}

template<class I>
__global__ __launch_bounds__(256) void gs_grid_fill_kernel(int64_t n, grid_desc g,
                                                           const I * __restrict__ rp,
                                                           const I * __restrict__ ci,
                                                           const double * __restrict__ vals,
                                                           double * __restrict__ gval,
                                                           uint8_t * __restrict__ gmask) {
	const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (r >= n) return;
	double v[7] = {0., 0., 0., 0., 0., 0., 0.};
	unsigned m = 0;
	for (I e = rp[r]; e < rp[r + 1]; ++e) {
		const int sl = grid_slot((int64_t)ci[e] - r, g);
		v[sl] = vals[e];
		m |= 1u << sl;
	}
#pragma unroll
	for (int k = 0; k < 7; ++k) gval[(int64_t)k * n + r] = v[k];
	gmask[r] = (uint8_t)m;
}

template<bool FWD>
__global__ __launch_bounds__(kGridThreads) void gs_tile_kernel(
	grid_desc g, const int32_t * __restrict__ tiles, int64_t first,
	const double * __restrict__ gval, const uint8_t * __restrict__ gmask, int64_t n,
	const double * __restrict__ b, double * x) {
	constexpr int T = kGridTile, H = kGridHalo;
	__shared__ double xs[kGridLds];
	const int tid = threadIdx.x;
	const int tile = tiles[first + blockIdx.x];
	const int64_t i0 = (int64_t)(tile % g.ti) * T;
	const int64_t j0 = (int64_t)((tile / g.ti) % g.tj) * T;
	const int64_t k0 = (int64_t)(tile / (g.ti * g.tj)) * T;
	// x on the tile and its halo (cells outside the grid: never read)
	for (int q = tid; q < kGridLds; q += kGridThreads) {
		const int64_t gi = i0 + q % H - 1, gj = j0 + (q / H) % H - 1, gk = k0 + q / (H * H) - 1;
		xs[q] = (gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny && gk >= 0 && gk < g.nz)
		            ? x[gi + g.nx * (gj + g.ny * gk)]
		            : 0.0;
	}
	const int li = tid % T, lj = (tid / T) % T, lk = tid / (T * T);
	const int64_t gi = i0 + li, gj = j0 + lj, gk = k0 + lk;
	const bool live = gi < g.nx && gj < g.ny && gk < g.nz;
	const int64_t r = live ? gi + g.nx * (gj + g.ny * gk) : 0;
	double v[7];
	unsigned m = 0;
	double br = 0.;
#pragma unroll
	for (int k = 0; k < 7; ++k) v[k] = 0.;
	if (live) {
#pragma unroll
		for (int k = 0; k < 7; ++k) v[k] = gval[(int64_t)k * n + r];
		m = gmask[r];
		br = b[r];
	}
	const int me = (li + 1) + H * ((lj + 1) + H * (lk + 1));
	const int mine = li + lj + lk;
	__syncthreads();
	for (int L = 0; L < 3 * T - 2; ++L) {
		if (live && mine == (FWD ? L : 3 * T - 3 - L)) {
			double rsum = 0.;
			if (m & 1u) rsum += v[0] * xs[me - H * H];
			if (m & 2u) rsum += v[1] * xs[me - H];
			if (m & 4u) rsum += v[2] * xs[me - 1];
			if (m & 16u) rsum += v[4] * xs[me + 1];
			if (m & 32u) rsum += v[5] * xs[me + H];
			if (m & 64u) rsum += v[6] * xs[me + H * H];
			const double dinv = 1. / v[3];
			xs[me] = dinv * (br - rsum);
		}
		__syncthreads();
	}
	if (live) x[r] = xs[me];
}

} // namespace fsx

using namespace fsx;

struct fsx_gs {
	fsx_ctx * ctx = nullptr;
	const fsx_csr * A = nullptr;
	int ncolours = 1;
	colour_map cm;
	// one sweep direction: level offsets, each level's ELL base, and the
	// level-ordered copy (row ids, entry counts, diagonal, tagged columns,
	// values) with the matrix entry each value came from
	struct dir {
		std::vector<int64_t> lev, base;
		std::vector<int32_t> width; // ELL depth of each level
		std::unique_ptr<dbuf> srow, slen, sdiag, scol, sval, dfrom, vfrom, dlev, dbase, dwidth;
		size_t nslots = 0;
	} fw, bw;
	// consecutive levels of at most this many rows (<= kRunThreads) run in
	// one workgroup (gs_run_kernel); 0: a launch per level
	int64_t run_rows = 0;
	std::unique_ptr<dbuf> xold;
	// stencil matrices on a grid (one colour): tiles swept by diagonal;
	// tile_ids ordered by diagonal, diagonal d's at [dptr[d], dptr[d + 1]),
	// the values by entry slot (gval) with their presence bits (gmask)
	bool tiles = false;
	grid_desc gd{};
	std::vector<int64_t> dptr;
	std::unique_ptr<dbuf> tile_ids, gval, gmask;
	uint64_t version = 0; // of the matrix values the copies hold
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

template<class T>
std::unique_ptr<dbuf> upload(fsx_ctx * ctx, const std::vector<T> & h) {
	auto d = std::make_unique<dbuf>(std::max<size_t>(1, h.size()) * sizeof(T));
	if (!h.empty())
		FSX_HIP(hipMemcpyAsync(d->p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice,
		                       ctx->stream));
	FSX_HIP(hipStreamSynchronize(ctx->stream)); // h may go out of scope
	return d;
}

// The grid of a one-colour stencil matrix, if it is one: every entry's
// offset c - r is 0 or +-1, +-nx, +-nx ny for a face neighbour of r's cell
// (no wrap-around), columns strictly ascending in each row, and the grid
// nx x ny x nz covers the rows.  The positive offsets give nx and nx ny
// (one offset: a line, two: a plane).
bool detect_grid(const std::vector<int64_t> & rp, const std::vector<int64_t> & ci,
                 grid_desc & g) {
	const int64_t n = (int64_t)rp.size() - 1;
	if (n < 2) return false;
	std::vector<int64_t> pos;
	for (int64_t r = 0; r < n && pos.size() <= 3; ++r)
		for (int64_t e = rp[r]; e < rp[r + 1]; ++e) {
			const int64_t o = ci[e] - r;
			if (o > 0 && std::find(pos.begin(), pos.end(), o) == pos.end()) pos.push_back(o);
		}
	if (pos.empty() || pos.size() > 3) return false;
	std::sort(pos.begin(), pos.end());
	if (pos[0] != 1) return false;
	g.nx = pos.size() >= 2 ? pos[1] : n;
	g.nxy = pos.size() == 3 ? pos[2] : n;
	if (g.nx < 2 || g.nxy % g.nx != 0 || n % g.nxy != 0) return false;
	if (pos.size() == 3 && g.nxy / g.nx < 2) return false;
	g.ny = g.nxy / g.nx;
	g.nz = n / g.nxy;
	const int64_t T = kGridTile;
	if ((g.nx + T - 1) / T * ((g.ny + T - 1) / T) * ((g.nz + T - 1) / T) >
	    std::numeric_limits<int32_t>::max())
		return false;
	g.ti = (int)((g.nx + T - 1) / T);
	g.tj = (int)((g.ny + T - 1) / T);
	g.tk = (int)((g.nz + T - 1) / T);
	std::atomic<bool> ok{true};
	for_row_blocks(n, setup_threads(n), [&](unsigned, int64_t lo, int64_t hi) {
		for (int64_t r = lo; r < hi && ok.load(std::memory_order_relaxed); ++r) {
			const int64_t i = r % g.nx, j = (r / g.nx) % g.ny, k = r / g.nxy;
			int64_t prev = -1;
			for (int64_t e = rp[r]; e < rp[r + 1]; ++e) {
				const int64_t c = ci[e], o = c - r;
				bool good = c > prev;
				if (o == 1) good = good && i + 1 < g.nx;
				else if (o == -1) good = good && i >= 1;
				else if (o == g.nx) good = good && j + 1 < g.ny;
				else if (o == -g.nx) good = good && j >= 1;
				else if (o == g.nxy) good = good && k + 1 < g.nz;
				else if (o == -g.nxy) good = good && k >= 1;
				else good = good && o == 0;
				if (!good) {
					ok = false;
					break;
				}
				prev = c;
			}
		}
	});
	return ok;
}

// Tiles ordered by diagonal I + J + K.
void build_tiles(fsx_ctx * ctx, fsx_gs & G) {
	const grid_desc & g = G.gd;
	const int nd = g.ti + g.tj + g.tk - 2;
	std::vector<int64_t> cnt((size_t)nd + 1, 0);
	for (int K = 0; K < g.tk; ++K)
		for (int J = 0; J < g.tj; ++J)
			for (int I = 0; I < g.ti; ++I) ++cnt[(size_t)(I + J + K) + 1];
	G.dptr.assign((size_t)nd + 1, 0);
	for (int d = 0; d < nd; ++d) G.dptr[(size_t)d + 1] = G.dptr[(size_t)d] + cnt[(size_t)d + 1];
	std::vector<int64_t> fill(G.dptr.begin(), G.dptr.end() - 1);
	std::vector<int32_t> ids((size_t)G.dptr.back());
	for (int K = 0; K < g.tk; ++K)
		for (int J = 0; J < g.tj; ++J)
			for (int I = 0; I < g.ti; ++I)
				ids[(size_t)fill[(size_t)(I + J + K)]++] = I + g.ti * (J + g.tj * K);
	G.tile_ids = upload(ctx, ids);
	const int64_t n = G.A->nrows;
	G.gval = std::make_unique<dbuf>(std::max<int64_t>(1, n) * 7 * sizeof(double));
	G.gmask = std::make_unique<dbuf>(std::max<int64_t>(1, n));
}

// The level-ordered copy of one sweep direction.  Entry order per row: the
// own colour's off-diagonal entries in CSR order, then the other colours'
// (gs.hh:94-107); the diagonal is the last entry whose column is the row.
template<class I>
void build_dir(fsx_ctx * ctx, const std::vector<int64_t> & rp, const std::vector<int64_t> & ci,
               const colour_map & cm, bool colours, bool fwd, fsx_gs::dir & D) {
	std::vector<int64_t> order;
	schedule(rp, ci, cm, colours, fwd, D.lev, order);
	const int64_t n = (int64_t)order.size();
	const size_t nl = D.lev.size() - 1;
	std::vector<I> srow(n);
	std::vector<int32_t> slen(n);
	std::vector<int64_t> dfrom(n, -1);
	D.base.assign(nl, 0);
	// levels are independent here: each level's width, then each level's
	// slots, on host threads taking levels in turn
	const unsigned nt = setup_threads(n);
	std::vector<int64_t> width(nl, 0);
	for_items(nl, nt, [&](size_t l) {
		int64_t w = 0;
		for (int64_t p = D.lev[l]; p < D.lev[l + 1]; ++p) {
			const int64_t r = order[p];
			int64_t len = 0;
			for (int64_t e = rp[r]; e < rp[r + 1]; ++e) len += ci[e] != r;
			w = std::max(w, len);
		}
		width[l] = w;
	});
	size_t nslots = 0;
	for (size_t l = 0; l < nl; ++l) {
		D.base[l] = (int64_t)nslots;
		nslots += (size_t)(width[l] * (D.lev[l + 1] - D.lev[l]));
	}
	std::vector<I> scol(nslots, (I)-1);
	std::vector<int64_t> vfrom(nslots, -1);
	for_items(nl, nt, [&](size_t l) {
		const int64_t cnt = D.lev[l + 1] - D.lev[l];
		for (int64_t p = D.lev[l]; p < D.lev[l + 1]; ++p) {
			const int64_t r = order[p], t = p - D.lev[l];
			const int64_t cr = colours ? cm.of(r) : 0;
			srow[p] = (I)r;
			int64_t j = 0;
			auto put = [&](int64_t e, bool own) {
				const int64_t c = ci[e];
				const bool visited = own && (fwd ? c < r : c > r);
				const int64_t slot = D.base[l] + j * cnt + t;
				scol[slot] = (I)(visited ? c : -c - 1);
				vfrom[slot] = e;
				++j;
			};
			for (int64_t e = rp[r]; e < rp[r + 1]; ++e) {
				const int64_t c = ci[e];
				if (colours && cm.of(c) != cr) continue;
				if (c == r) dfrom[p] = e;
				else put(e, true);
			}
			if (colours)
				for (int64_t e = rp[r]; e < rp[r + 1]; ++e)
					if (cm.of(ci[e]) != cr) put(e, false);
			slen[p] = (int32_t)j;
		}
	});
	D.nslots = nslots;
	D.width.resize(nl);
	for (size_t l = 0; l < nl; ++l) {
		FSX_REQUIRE(width[l] <= std::numeric_limits<int32_t>::max(), "hybrid_gs: row too long");
		D.width[l] = (int32_t)width[l];
	}
	D.dlev = upload(ctx, D.lev);
	D.dbase = upload(ctx, D.base);
	D.dwidth = upload(ctx, D.width);
	D.srow = upload(ctx, srow);
	D.slen = upload(ctx, slen);
	D.scol = upload(ctx, scol);
	D.dfrom = upload(ctx, dfrom);
	D.vfrom = upload(ctx, vfrom);
	D.sdiag = std::make_unique<dbuf>(std::max<int64_t>(1, n) * sizeof(double));
	D.sval = std::make_unique<dbuf>(std::max<size_t>(1, nslots) * sizeof(double));
}

void refresh_values(const fsx_gs & G) {
	hipStream_t s = G.ctx->stream;
	const double * vals = (const double *)G.A->vals->p;
	if (G.tiles) {
		const int64_t n = G.A->nrows;
		const dim3 grid((unsigned)std::max<int64_t>(1, (n + 255) / 256));
		if (G.A->it == FSX_I32)
			hipLaunchKernelGGL((gs_grid_fill_kernel<int32_t>), grid, dim3(256), 0, s, n, G.gd,
			                   (const int32_t *)G.A->rowptr(), (const int32_t *)G.A->colind(),
			                   vals, (double *)G.gval->p, (uint8_t *)G.gmask->p);
		else
			hipLaunchKernelGGL((gs_grid_fill_kernel<int64_t>), grid, dim3(256), 0, s, n, G.gd,
			                   (const int64_t *)G.A->rowptr(), (const int64_t *)G.A->colind(),
			                   vals, (double *)G.gval->p, (uint8_t *)G.gmask->p);
		FSX_HIP(hipGetLastError());
		return;
	}
	for (const fsx_gs::dir * D : {&G.fw, &G.bw}) {
		const size_t n = (size_t)G.A->nrows;
		hipLaunchKernelGGL(gs_gather_kernel, dim3((unsigned)std::max<size_t>(1, (n + 255) / 256)),
		                   dim3(256), 0, s, n, (const int64_t *)D->dfrom->p, vals,
		                   (double *)D->sdiag->p);
		if (D->nslots)
			hipLaunchKernelGGL(gs_gather_kernel, dim3((unsigned)((D->nslots + 255) / 256)),
			                   dim3(256), 0, s, D->nslots, (const int64_t *)D->vfrom->p, vals,
			                   (double *)D->sval->p);
	}
	FSX_HIP(hipGetLastError());
}

template<class I>
void enqueue_sweep(const fsx_gs & G, bool fwd, const double * b, double * x) {
	const fsx_csr * A = G.A;
	hipStream_t s = G.ctx->stream;
	const size_t n = (size_t)A->nrows;
	if (G.tiles) {
		// one launch per tile diagonal, in sweep order
		const int nd = (int)G.dptr.size() - 1;
		for (int q = 0; q < nd; ++q) {
			const int d = fwd ? q : nd - 1 - q;
			const int64_t first = G.dptr[(size_t)d];
			const unsigned cnt = (unsigned)(G.dptr[(size_t)d + 1] - first);
			if (fwd)
				hipLaunchKernelGGL((gs_tile_kernel<true>), dim3(cnt), dim3(kGridThreads), 0, s,
				                   G.gd, (const int32_t *)G.tile_ids->p, first,
				                   (const double *)G.gval->p, (const uint8_t *)G.gmask->p,
				                   (int64_t)n, b, x);
			else
				hipLaunchKernelGGL((gs_tile_kernel<false>), dim3(cnt), dim3(kGridThreads), 0, s,
				                   G.gd, (const int32_t *)G.tile_ids->p, first,
				                   (const double *)G.gval->p, (const uint8_t *)G.gmask->p,
				                   (int64_t)n, b, x);
		}
		FSX_HIP(hipGetLastError());
		return;
	}
	hipLaunchKernelGGL(gs_snapshot_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
	                   n, (const double *)x, (double *)G.xold->p);
	const fsx_gs::dir & D = fwd ? G.fw : G.bw;
	const double * xo = (const double *)G.xold->p;
	const size_t nl = D.lev.size() - 1;
	auto small = [&](size_t l) {
		return G.run_rows > 0 && D.lev[l + 1] - D.lev[l] <= G.run_rows;
	};
	for (size_t l = 0; l < nl;) {
		size_t e = l + 1; // a run: consecutive small levels, at most kRunMaxLevels
		if (small(l))
			while (e < nl && e - l < (size_t)kRunMaxLevels && small(e)) ++e;
		if (e - l >= 2) {
			hipLaunchKernelGGL((gs_run_kernel<I>), dim3(1), dim3(kRunThreads), 0, s,
			                   (const I *)D.srow->p, (const int32_t *)D.slen->p,
			                   (const double *)D.sdiag->p, (const I *)D.scol->p,
			                   (const double *)D.sval->p, (const int64_t *)D.dlev->p,
			                   (const int64_t *)D.dbase->p, (const int32_t *)D.dwidth->p, (int64_t)l,
			                   (int)(e - l), b, x, xo);
			l = e;
			continue;
		}
		const int64_t cnt = D.lev[l + 1] - D.lev[l];
		const unsigned g = (unsigned)((cnt + 255) / 256);
		hipLaunchKernelGGL((gs_level_kernel<I>), dim3(g), dim3(256), 0, s,
		                   (const I *)D.srow->p, (const int32_t *)D.slen->p,
		                   (const double *)D.sdiag->p, (const I *)D.scol->p,
		                   (const double *)D.sval->p, D.lev[l], cnt, D.base[l], D.width[l], b, x,
		                   xo);
		++l;
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
			for_row_blocks(A->nnz, setup_threads(A->nnz), [&](unsigned, int64_t lo, int64_t hi) {
				for (int64_t e = lo; e < hi; ++e) {
					ci[e] = isz == 4 ? ((int32_t *)hci.data())[e] : ((int64_t *)hci.data())[e];
					FSX_REQUIRE(ci[e] >= 0 && ci[e] < n, "hybrid_gs: column index out of range");
				}
			});
		}
		auto G = std::make_unique<fsx_gs>();
		G->ctx = ctx;
		G->A = A;
		G->ncolours = ncolours;
		if (n > 0) {
			G->cm.q = n / ncolours;
			G->cm.m = n % ncolours;
			G->cm.big = G->cm.m * (G->cm.q + 1);
		}
		// one colour on a grid: tiles (FSX_GS_TILES=0 keeps the level copies)
		const char * te = getenv("FSX_GS_TILES");
		G->tiles = ncolours == 1 && !(te && te[0] == '0') && detect_grid(rp, ci, G->gd);
		const char * re = getenv("FSX_GS_RUN_ROWS");
		G->run_rows = std::min<int64_t>(kRunThreads,
		                                re ? std::max<int64_t>(0, atoll(re)) : kRunRowsDefault);
		for (bool fwd : {true, false}) {
			fsx_gs::dir & D = fwd ? G->fw : G->bw;
			if (G->tiles) { // the level count only (fsx_gs_levels)
				std::vector<int64_t> order;
				schedule(rp, ci, G->cm, false, fwd, D.lev, order);
			}
			else if (isz == 4) build_dir<int32_t>(ctx, rp, ci, G->cm, ncolours > 1, fwd, D);
			else build_dir<int64_t>(ctx, rp, ci, G->cm, ncolours > 1, fwd, D);
		}
		if (G->tiles) build_tiles(ctx, *G);
		G->xold = std::make_unique<dbuf>(std::max<int64_t>(1, n) * sizeof(double));
		refresh_values(*G);
		G->version = A->vals_version;
		ctx_retain(ctx);
		*out = G.release();
	});
}

int fsx_gs_levels(const fsx_gs * G, int64_t * forward, int64_t * backward) {
	return guard([&] {
		FSX_REQUIRE(G, "null smoother");
		if (forward) *forward = (int64_t)G->fw.lev.size() - 1;
		if (backward) *backward = (int64_t)G->bw.lev.size() - 1;
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
		if (G->version != G->A->vals_version) { // values changed in place
			refresh_values(*G);
			G->version = G->A->vals_version;
		}
		const double * bp = (const double *)b->ptr();
		double * xp = (double *)x->ptr();
		if (!gs_graphs()) {
			enqueue_relax(*G, sweep, nrelax, bp, xp);
			return;
		}
		std::vector<const void *> key{bp, xp, (const void *)(intptr_t)sweep,
		                              (const void *)(intptr_t)nrelax};
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
