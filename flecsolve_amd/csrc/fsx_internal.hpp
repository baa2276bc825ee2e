// fsx_internal.hpp -- internal types shared by the libfsx translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "fsx.h"

namespace fsx {

// ---- error plumbing ------------------------------------------------------

void set_error(const std::string & msg);
const char * get_error();

struct error : std::exception {
	int code;
	std::string msg;
	error(int c, std::string m) : code(c), msg(std::move(m)) {}
	const char * what() const noexcept override { return msg.c_str(); }
};

#define FSX_HIP(call)                                                        \
	do {                                                                     \
		hipError_t e_ = (call);                                              \
		if (e_ != hipSuccess)                                                \
			throw ::fsx::error(FSX_ERR_HIP,                                  \
			                   std::string(#call ": ") + hipGetErrorString(e_)); \
	} while (0)

#define FSX_REQUIRE(cond, msg)                                               \
	do {                                                                     \
		if (!(cond)) throw ::fsx::error(FSX_ERR_ARG, (msg));                 \
	} while (0)

// Run `body` and translate exceptions into status codes for the C ABI.
template<class F>
int guard(F && body) {
	try {
		body();
		return FSX_OK;
	}
	catch (const error & e) {
		set_error(e.msg);
		return e.code;
	}
	catch (const std::bad_alloc &) {
		set_error("host allocation failed");
		return FSX_ERR_ALLOC;
	}
	catch (const std::exception & e) {
		set_error(e.what());
		return FSX_ERR_STATE;
	}
	catch (...) {
		set_error("unknown error");
		return FSX_ERR_STATE;
	}
}

// ---- device buffers ------------------------------------------------------

// Live device-buffer ledger: every dbuf records its address range while it
// lives (allocations and wrapped caller memory alike).  Cached HIP graphs
// keep the ranges their launches touch and are checked against it before
// every replay, so a graph whose captured addresses are no longer backed by
// a live buffer of at least the captured size fails with an error instead of
// launching.
namespace ledger {
struct span {
	const void * p;
	size_t bytes;
};
void add(const void * p, size_t bytes);
void remove(const void * p, size_t bytes);
// true when [p, p + bytes) lies inside one live buffer
bool live(const void * p, size_t bytes);
} // namespace ledger

struct dbuf {
	void * p = nullptr;
	size_t bytes = 0;
	bool owns = true;
	dbuf() = default;
	explicit dbuf(size_t b);
	~dbuf();
	// refer to caller-owned device memory (not freed here)
	void adopt(void * q, size_t b);
	dbuf(const dbuf &) = delete;
	dbuf & operator=(const dbuf &) = delete;
};

// Reduction workspace: per-block partials and a small array of device
// scalars.
constexpr int kRedBlocks = 2048;   // maximum blocks of a reduction grid
constexpr int kScalarSlots = 64;   // device scalar slots per context
constexpr int kStencilMax = 8;     // offsets of the symmetric stencil layout

struct reduce_ws {
	double * partials = nullptr;   // grown on demand (ensure_partials)
	double * scalars = nullptr;    // [kScalarSlots]
};

} // namespace fsx

struct cg_graph_cache;
struct cg_cache_deleter {
	void operator()(cg_graph_cache * c) const; // defined in fsx_cg.hip
};

struct fsx_ctx {
	// Vectors and matrices hold a reference; the context is freed when the
	// owner's fsx_ctx_destroy and the last dependent release have happened.
	std::atomic<int> refs{1};
	int device = 0;
	hipStream_t stream = nullptr;
	int num_cus = 0;
	std::unique_ptr<fsx::dbuf> ws_buf;
	std::unique_ptr<fsx::dbuf> part_buf; // per-block partials
	size_t part_cap = 0;                 // in doubles
	std::unique_ptr<fsx::dbuf> relax_buf; // relaxation sweep buffer
	size_t relax_cap = 0;                 // in doubles
	fsx::reduce_ws ws;
	double * host_scalars = nullptr; // pinned, kScalarSlots
	uint64_t host_syncs = 0;
	uint64_t launches = 0;
	uint64_t next_id = 1;
	std::vector<double> last_cg_hist;
	std::unique_ptr<cg_graph_cache, cg_cache_deleter> cg_cache;
	~fsx_ctx();
};

struct fsx_vec {
	std::atomic<int> refcnt{1};
	fsx_ctx * ctx = nullptr;
	size_t n = 0;
	fsx_dtype dt = FSX_F64;
	std::shared_ptr<fsx::dbuf> buf;
	uint64_t id = 0;
	void * ptr() const { return buf->p; }
	size_t elem() const { return dt == FSX_F64 ? 8 : 4; }
};

struct csr_index {
	std::unique_ptr<fsx::dbuf> rowptr, colind;
	// Sliced-row layout (built for short-row matrices): rows grouped in
	// slices of 64 (one wavefront), each slice padded to its longest row and
	// stored entry-major, so lane l of a wavefront reads entry j of row
	// 64*s + l at sl_ptr[s] + 64*j + l.  Entries keep their CSR order within
	// each row; padding carries column -1.  Shared by all value copies.
	std::unique_ptr<fsx::dbuf> sl_ptr, sl_col;
	int64_t sl_nnz = 0;
	// Slice-offset layout (built for matrices whose 64-row slices share a
	// few column offsets, e.g. stencils): slice s keeps the sorted offsets
	// off[q] = column - row present in any of its rows (q = dg_ptr[s]/64 +
	// j), a 64-bit mask of the lanes whose row has an entry at offset j, and
	// values entry-major at dg_ptr[s] + 64 j + lane (zero where absent).
	// Index bytes per entry: 12/64 instead of 4; rows keep CSR order.
	std::unique_ptr<fsx::dbuf> dg_ptr, dg_off, dg_mask;
	int64_t dg_nnz = 0;
	// Symmetric stencil layout (built next to the slice-offset one when every
	// row's offsets come from one ascending list of at most kStencilMax
	// offsets and the structure is symmetric): st_off lists the offsets; a
	// 64-bit lane mask per (slice, offset) at st_mask[s * st_w + j]; values
	// only for the st_wu offsets >= 0, entry-major per slice at
	// (s * st_wu + st_ju[j]) * 64 + lane.  The entry of row r at offset
	// o < 0 is read from row r + o's stored entry at -o (st_ju[j] names that
	// column), which holds the same value when the matrix is symmetric.
	// st_state: -1 not analysed, 0 not applicable, 1 built.
	int st_state = -1;
	int st_w = 0, st_wu = 0;
	int32_t st_off[fsx::kStencilMax] = {};
	int32_t st_ju[fsx::kStencilMax] = {};
	std::unique_ptr<fsx::dbuf> st_mask;
	// Lazily built transpose structure for A^T x (adjoint).
	std::unique_ptr<fsx::dbuf> t_rowptr, t_colind, t_perm;
	// every row stores its diagonal entry: -1 not yet checked, 0 no, 1 yes;
	// diag_pos (int64 per row, -1 when absent) once checked; diag_bits: one
	// bit per entry, set at each diag_pos entry
	int diag_full = -1;
	std::unique_ptr<fsx::dbuf> diag_pos, diag_bits;
};

struct fsx_csr {
	fsx_ctx * ctx = nullptr;
	int64_t nrows = 0, ncols = 0, nnz = 0;
	fsx_itype it = FSX_I32;
	fsx_dtype vt = FSX_F64;
	std::shared_ptr<csr_index> idx;
	std::unique_ptr<fsx::dbuf> vals;
	std::unique_ptr<fsx::dbuf> sl_vals; // values in the sliced-row layout
	std::unique_ptr<fsx::dbuf> dg_vals; // values in the slice-offset layout
	// values in the symmetric stencil layout; st_sym: the CSR values are
	// bitwise symmetric (1), not (0), or not yet checked (-1)
	std::unique_ptr<fsx::dbuf> st_vals;
	int st_sym = -1;
	// dg_vals lag the CSR values (a rewrite served by the stencil layout
	// skipped them); rebuilt by ensure_dg_fresh before any use
	bool dg_stale = false;
	// SpMV row mapping: 0 LDS-staged row blocks, 1 vector per row, 2 sliced
	// rows, 3 slice offsets, 4 symmetric stencil
	int plan_kind = 0;
	int plan_lanes = 0;
	int64_t max_row = 0;
	double avg_row = 0;
	// bumped whenever the values change in place (copies built from them,
	// e.g. a smoother's, check it)
	uint64_t vals_version = 0;
	const void * rowptr() const { return idx->rowptr->p; }
	const void * colind() const { return idx->colind->p; }
};

namespace fsx {

// ---- helpers implemented in the .hip files -------------------------------

void ctx_retain(fsx_ctx * c);
void ctx_release(fsx_ctx * c);
// Rebuild A's slice-offset values from its CSR values if they lag
// (stream-ordered; A's layouts are caches of its CSR values).
void ensure_dg_fresh(fsx_ctx * ctx, const fsx_csr * A);

inline fsx_ctx * ctx_of(const fsx_vec * v) { return v->ctx; }
void check_vec(const fsx_vec * v, const char * name);
void check_same(const fsx_vec * a, const fsx_vec * b);

// Reductions into a device scalar slot (async) and to the host (sync).
enum red_kind { RED_DOT = 0, RED_L1, RED_L2SQ, RED_INF, RED_MIN, RED_MAX };
void reduce_async(fsx_ctx * ctx, red_kind k, const fsx_vec * x,
                  const fsx_vec * y, double * dslot);
double reduce_host(fsx_ctx * ctx, red_kind k, const fsx_vec * x,
                   const fsx_vec * y);
// The same reductions over n elements of dtype dt at x (and y for a dot),
// with the caller's partials (at least red_grid(n) doubles).  Allocates
// nothing.
void reduce_raw(fsx_ctx * ctx, red_kind k, fsx_dtype dt, size_t n, const void * x,
                const void * y, double * dslot, double * part);

// s_r = 1 / sqrt(a_rr) of the FP64 matrix A on ctx's stream; *bad (device,
// optional) is set to 1 when a diagonal entry is missing or not positive.
void sym_scale_vector(fsx_ctx * ctx, const fsx_csr * A, double * s, int * bad);

// Elementwise kernels used by the solve drivers.
void launch_axpy(fsx_ctx * ctx, double a, const fsx_vec * x, const fsx_vec * y,
                 fsx_vec * z);
void launch_copy(fsx_ctx * ctx, const fsx_vec * s, fsx_vec * d);
void launch_set(fsx_ctx * ctx, double a, fsx_vec * x);
void launch_convert(fsx_ctx * ctx, const fsx_vec * s, fsx_vec * d, double scale);

// SpMV row-mapping selection (fsx_spmv.hip); may build the sliced or the
// slice-offset layout from the host copies of the structure
void choose_plan(fsx_ctx * ctx, fsx_csr * A, const void * host_rowptr,
                 const void * host_colind);
// Build the slice-offset layout if A's structure qualifies (and, with
// check_budget, its padding is small enough); false otherwise.
bool ensure_sdia(fsx_ctx * ctx, fsx_csr * A, const void * host_rowptr,
                 const void * host_colind, bool check_budget);
// Build (or complete) the sliced-row layout of A; host_rowptr may be null.
void ensure_sliced(fsx_ctx * ctx, fsx_csr * A, const void * host_rowptr);
// After A->vals changed on the device: refill the sliced and slice-offset
// copies in place.
void refresh_sliced_values(fsx_ctx * ctx, fsx_csr * A);

// Make ctx->ws.partials hold at least `count` doubles.  Reallocation drops
// captured graphs (they hold the old address).
double * ensure_partials(fsx_ctx * ctx, size_t count);
// A context-owned vector of at least n doubles for relaxation sweeps (not
// used inside captured graphs; reallocation synchronises the stream).
double * ensure_relax_buffer(fsx_ctx * ctx, size_t n);
// Drop captured solver graphs of ctx (defined with the CG driver).
void invalidate_graphs(fsx_ctx * ctx);
// Look up the no-allocation entry point's scratch array on `device` (the
// current device), once per device; called by fsx_ctx_create.
void noalloc_init(int device);

// host-side scalar fetch of a device slot (sync)
double fetch_scalar(fsx_ctx * ctx, const double * dslot);

// grid size for a reduction over n elements (fixed function of n)
int red_grid(size_t n);

template<class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;
}

// Sum over a 256-thread block; result valid in thread 0.  `sh` holds 4 values.
template<class T>
__device__ __forceinline__ T block_sum_256(T v, T * sh) {
	v = wave_sum(v);
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (lane == 0) sh[w] = v;
	__syncthreads();
	T r = T(0);
	if (threadIdx.x == 0) r = ((sh[0] + sh[1]) + sh[2]) + sh[3];
	return r;
}

// Two-stage deterministic reductions.  Stage 1: every block reduces its
// terms and thread 0 stores them at partials[blk*NV + i] with plain stores
// (no fences or atomics: on a multi-XCD GPU a device-scope release fence per
// block would write back the XCD's L2 each time).  Stage 2: a one-block
// kernel, launched next in stream order (the kernel boundary makes the
// partials visible), sums them in a fixed order.
// Blocks of up to 1024 threads (16 waves); the waves' sums are added in
// wave order.
template<int NV>
__device__ __forceinline__ void store_block_partials(const double (&d)[NV],
                                                     double * partials,
                                                     unsigned blk) {
	__shared__ double sh[NV][16];
#pragma unroll
	for (int i = 0; i < NV; ++i) {
		double s = wave_sum(d[i]);
		if ((threadIdx.x & 63) == 0) sh[i][threadIdx.x >> 6] = s;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		const int nw = (int)(blockDim.x >> 6);
#pragma unroll
		for (int i = 0; i < NV; ++i) {
			double t = sh[i][0];
			for (int q = 1; q < nw; ++q) t += sh[i][q];
			partials[(size_t)blk * NV + i] = t;
		}
	}
}

constexpr int kFinThreads = 1024;

// Sum of nblk*NV partials by one block of NT threads (kFinThreads unless
// stated).  Thread t owns blocks t, t + NT, ...; it keeps kFinUnroll
// independent accumulators so that many loads are in flight (the reads are
// latency bound), folds them in a fixed order, then a wave tree and the
// waves in order.  The order depends only on nblk and NT: deterministic, so
// every block that sums the same partials gets the same bits.  Result valid
// in every thread.
constexpr int kFinUnroll = 8;

template<int NV, int NT = kFinThreads>
__device__ __forceinline__ void sum_partials(const double * partials,
                                             unsigned nblk, double (&tot)[NV]) {
	__shared__ double red[NV][NT / 64];
	double acc[NV][kFinUnroll];
#pragma unroll
	for (int i = 0; i < NV; ++i)
#pragma unroll
		for (int u = 0; u < kFinUnroll; ++u) acc[i][u] = 0.0;
	unsigned b = threadIdx.x;
	for (; b + (kFinUnroll - 1) * NT < nblk; b += kFinUnroll * NT) {
#pragma unroll
		for (int u = 0; u < kFinUnroll; ++u)
#pragma unroll
			for (int i = 0; i < NV; ++i)
				acc[i][u] += partials[(size_t)(b + u * NT) * NV + i];
	}
	for (; b < nblk; b += NT)
#pragma unroll
		for (int i = 0; i < NV; ++i) acc[i][0] += partials[(size_t)b * NV + i];
#pragma unroll
	for (int i = 0; i < NV; ++i) {
		double a = 0.0;
#pragma unroll
		for (int u = 0; u < kFinUnroll; ++u) a += acc[i][u];
		double s = wave_sum(a);
		if ((threadIdx.x & 63) == 0) red[i][threadIdx.x >> 6] = s;
	}
	__syncthreads();
#pragma unroll
	for (int i = 0; i < NV; ++i) {
		double s = 0.0;
		for (int w = 0; w < NT / 64; ++w) s += red[i][w];
		tot[i] = s;
	}
}

// Stage-2 kernel: f(tot) runs on thread 0 with the grid totals.
template<int NV, class F>
__global__ __launch_bounds__(kFinThreads) void finalize_kernel(
	const double * partials, unsigned nblk, F f) {
	if (f.skip()) return;
	double tot[NV];
	sum_partials<NV>(partials, nblk, tot);
	if (threadIdx.x == 0) f(tot);
}

// Finaliser that stores the totals (optionally square-rooted) to `out`.
template<int NV>
struct fin_store {
	double * out;
	__device__ bool skip() const { return false; }
	__device__ void operator()(const double (&t)[NV]) const {
#pragma unroll
		for (int i = 0; i < NV; ++i) out[i] = t[i];
	}
};

} // namespace fsx
