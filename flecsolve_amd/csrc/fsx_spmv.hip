// fsx_spmv.hip -- device CSR operator: creation, SpMV, fused residual and
// SpMV+dot, Jacobi relaxation, transpose product and timing helpers.
//
// Replaces mat::parcsr's apply (flecsolve/matrices/parcsr.hh:60-91): at one
// colour the off-diagonal block is empty, the diagonal block is the whole
// matrix with unchanged column ids (topo/csr.hh:553-618), and the product is
// compressed_ops::spmv (matrices/seq.hh:177-193) -- here one kernel.
#include "fsx_spmv_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace fsx {

// ---- epilogues -------------------------------------------------------------

template<class ACC, class Y>
struct epi_store {
	static constexpr int NV = 1;
	static constexpr bool reduces = false;
	Y * y;
	__device__ bool skip() const { return false; }
	template<class XI>
	__device__ void row(int64_t i, ACC s, double (&)[NV], XI &&) const { y[i] = (Y)s; }
	__device__ void block_end(const double (&)[NV], unsigned, unsigned) const {}
};

// r = b - A x  (op::core::residual: apply, then subtract_self<false>,
// operators/core.hh:135-142 and topo_tasks.hh:98-108)
template<class ACC, class Y>
struct epi_residual {
	static constexpr int NV = 1;
	static constexpr bool reduces = false;
	Y * r;
	const Y * b;
	__device__ bool skip() const { return false; }
	template<class XI>
	__device__ void row(int64_t i, ACC s, double (&)[NV], XI &&) const {
		r[i] = b[i] - (Y)s;
	}
	__device__ void block_end(const double (&)[NV], unsigned, unsigned) const {}
};

// One weighted Jacobi sweep from xin (the x of the sweep's start) into
// xout: x_i = omega (1/a_ii)(b_i - sum_{j != i} a_ij xin_j) + (1 - omega)
// xin_i, the off-diagonal sum in CSR order (mg::bound_jacobi::relax,
// solvers/mg/jacobi.hh:58-93).  Served by the sliced and slice-offset
// kernels only.
struct epi_relax {
	static constexpr int NV = 1;
	static constexpr bool reduces = false;
	static constexpr bool offdiag = true;
	double omega;
	const double * b;
	double * xout;
	__device__ bool skip() const { return false; }
	template<class XI>
	__device__ void row_offdiag(int64_t i, double s, double diag, XI && xi) const {
		const double dinv = 1. / diag;
		xout[i] = omega * dinv * (b[i] - s) + (1 - omega) * xi();
	}
	template<class XI>
	__device__ void row(int64_t, double, double (&)[NV], XI &&) const {}
	__device__ void block_end(const double (&)[NV], unsigned, unsigned) const {}
};

// w = A p and the per-block partials of w . p  (cg.hh:95-98: A.apply(p, w);
// w.dot(p)); a finalize kernel sums the partials.
template<class ACC, class Y>
struct epi_dot {
	static constexpr int NV = 1;
	static constexpr bool reduces = true;
	Y * w;
	const Y * p;
	double * partials;
	__device__ bool skip() const { return false; }
	template<class XI>
	__device__ void row(int64_t i, ACC s, double (&d)[NV], XI && pi) const {
		const double pv = (double)pi(); // load before the store
		const Y wi = (Y)s;
		w[i] = wi;
		d[0] += (double)wi * pv;
	}
	__device__ void block_end(const double (&d)[NV], unsigned blk,
	                          unsigned) const {
		store_block_partials<NV>(d, partials, blk);
	}
};

// ---- plan selection and the sliced-row layout ---------------------------

// Scatter CSR rows into the sliced layout (one thread per slot row, padding
// rows of the last slice included).  With cols == nullptr only values.
template<class I, class V>
__global__ void sliced_build_kernel(int64_t nrows, int64_t nslots,
                                    const I * __restrict__ rp,
                                    const I * __restrict__ ci,
                                    const V * __restrict__ v,
                                    const I * __restrict__ sl_ptr,
                                    I * __restrict__ sl_col,
                                    V * __restrict__ sl_val) {
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= nslots) return;
	const int64_t s = r / kSliceRows;
	const int lane = (int)(r % kSliceRows);
	const I base = sl_ptr[s];
	const int w = (int)((sl_ptr[s + 1] - base) / kSliceRows);
	int64_t lo = 0, len = 0;
	if (r < nrows) {
		lo = rp[r];
		len = rp[r + 1] - lo;
	}
	for (int j = 0; j < w; ++j) {
		const int64_t e = (int64_t)base + (int64_t)j * kSliceRows + lane;
		if (j < len) {
			if (sl_col) sl_col[e] = ci[lo + j];
			sl_val[e] = v[lo + j];
		}
		else {
			if (sl_col) sl_col[e] = (I)-1;
			sl_val[e] = V(0);
		}
	}
}

// Slice offsets (64-bit) from a host row-pointer array of index type I.
template<class I>
static std::vector<int64_t> slice_offsets(int64_t nrows, const I * hrp) {
	const int64_t ns = (nrows + kSliceRows - 1) / kSliceRows;
	std::vector<int64_t> sp(ns + 1);
	int64_t tot = 0;
	for (int64_t s = 0; s < ns; ++s) {
		sp[s] = tot;
		int64_t w = 0;
		const int64_t hi = std::min<int64_t>(nrows, (s + 1) * kSliceRows);
		for (int64_t r = s * kSliceRows; r < hi; ++r)
			w = std::max<int64_t>(w, (int64_t)hrp[r + 1] - hrp[r]);
		tot += w * kSliceRows;
	}
	sp[ns] = tot;
	return sp;
}

template<class I>
static void build_sliced_t(fsx_ctx * ctx, fsx_csr * A, const void * host_rp) {
	auto & X = *A->idx;
	const int64_t ns = (A->nrows + kSliceRows - 1) / kSliceRows;
	const int64_t nslots = ns * kSliceRows;
	const bool need_index = !X.sl_ptr;
	if (need_index) {
		std::vector<I> hrp;
		const I * rp = (const I *)host_rp;
		if (!rp) {
			hrp.resize(A->nrows + 1);
			FSX_HIP(hipMemcpyAsync(hrp.data(), A->rowptr(), hrp.size() * sizeof(I),
			                       hipMemcpyDeviceToHost, ctx->stream));
			FSX_HIP(hipStreamSynchronize(ctx->stream));
			rp = hrp.data();
		}
		std::vector<int64_t> sp64 = slice_offsets<I>(A->nrows, rp);
		FSX_REQUIRE(sizeof(I) == 8 || sp64[ns] < ((int64_t)1 << 31),
		            "sliced layout not addressable with int32 indices");
		std::vector<I> sp(sp64.begin(), sp64.end());
		X.sl_nnz = sp64[ns];
		X.sl_ptr = std::make_unique<dbuf>(sp.size() * sizeof(I));
		X.sl_col = std::make_unique<dbuf>(std::max<int64_t>(1, X.sl_nnz) * sizeof(I));
		FSX_HIP(hipMemcpyAsync(X.sl_ptr->p, sp.data(), sp.size() * sizeof(I),
		                       hipMemcpyHostToDevice, ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
	}
	const size_t vsz = A->vt == FSX_F64 ? 8 : 4;
	A->sl_vals = std::make_unique<dbuf>(std::max<int64_t>(1, X.sl_nnz) * vsz);
	if (nslots == 0) return;
	const unsigned g = (unsigned)((nslots + 255) / 256);
	I * col = need_index ? (I *)X.sl_col->p : nullptr;
	if (A->vt == FSX_F64)
		hipLaunchKernelGGL((sliced_build_kernel<I, double>), dim3(g), dim3(256), 0,
		                   ctx->stream, A->nrows, nslots, (const I *)A->rowptr(),
		                   (const I *)A->colind(), (const double *)A->vals->p,
		                   (const I *)X.sl_ptr->p, col, (double *)A->sl_vals->p);
	else
		hipLaunchKernelGGL((sliced_build_kernel<I, float>), dim3(g), dim3(256), 0,
		                   ctx->stream, A->nrows, nslots, (const I *)A->rowptr(),
		                   (const I *)A->colind(), (const float *)A->vals->p,
		                   (const I *)X.sl_ptr->p, col, (float *)A->sl_vals->p);
	FSX_HIP(hipGetLastError());
	FSX_HIP(hipStreamSynchronize(ctx->stream));
}

void ensure_sliced(fsx_ctx * ctx, fsx_csr * A, const void * host_rowptr) {
	if (A->idx->sl_ptr && A->sl_vals) return;
	if (A->it == FSX_I32) build_sliced_t<int32_t>(ctx, A, host_rowptr);
	else build_sliced_t<int64_t>(ctx, A, host_rowptr);
}

// Re-copy A's CSR values into its existing sliced-row buffer (same address,
// so captured graphs stay valid); stream-ordered, no sync.
template<class I, class V>
static void refresh_sliced_t(fsx_ctx * ctx, fsx_csr * A) {
	const int64_t nslots = ((A->nrows + kSliceRows - 1) / kSliceRows) * kSliceRows;
	if (nslots == 0) return;
	const unsigned g = (unsigned)((nslots + 255) / 256);
	hipLaunchKernelGGL((sliced_build_kernel<I, V>), dim3(g), dim3(256), 0, ctx->stream,
	                   A->nrows, nslots, (const I *)A->rowptr(), (const I *)A->colind(),
	                   (const V *)A->vals->p, (const I *)A->idx->sl_ptr->p, (I *)nullptr,
	                   (V *)A->sl_vals->p);
	FSX_HIP(hipGetLastError());
	ctx->launches++;
}

// ---- slice-offset layout ----------------------------------------------------

// Values of row r (thread r of the padded slice range) into the layout:
// zero the row's slots, then place each entry at its offset's position.
// Rows and offsets are both ascending, so one forward scan finds them.
template<class I, class V>
__global__ void sdia_scatter_kernel(int64_t nrows, int64_t nslots,
                                    const I * __restrict__ rp, const I * __restrict__ ci,
                                    const V * __restrict__ v,
                                    const int64_t * __restrict__ dg_ptr,
                                    const int32_t * __restrict__ dg_off,
                                    V * __restrict__ dg_val) {
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= nslots) return;
	const int64_t s = r / kSliceRows;
	const int lane = (int)(r % kSliceRows);
	const int64_t base = dg_ptr[s];
	const int w = (int)((dg_ptr[s + 1] - base) / kSliceRows);
	const int64_t q0 = base / kSliceRows;
	for (int j = 0; j < w; ++j) dg_val[base + (int64_t)j * kSliceRows + lane] = V(0);
	if (r >= nrows) return;
	int j = 0;
	for (I e = rp[r]; e < rp[r + 1]; ++e) {
		const int64_t off = (int64_t)ci[e] - r;
		while (j < w && (int64_t)dg_off[q0 + j] != off) ++j;
		if (j == w) return; // not this layout's structure (cannot happen)
		dg_val[base + (int64_t)j * kSliceRows + lane] = v[e];
		++j;
	}
}

// Host analysis: per slice, the sorted union of the rows' column offsets and
// the lane masks.  Qualifies when every row's columns strictly increase, no
// slice has more than 64 offsets and offsets fit int32; for automatic
// selection also the padded size must be within 25% of nnz (+ one slice
// column): a padded slot costs one value
// (8 or 4 bytes) while every entry saves its 4-byte column id, so the layout
// moves fewer bytes than the sliced one up to ~50% (FP64) / ~100% (FP32)
// padding.
template<class I>
static bool sdia_analyze(int64_t nrows, int64_t nnz, const I * rp, const I * ci,
                         bool check_budget, std::vector<int64_t> & ptr,
                         std::vector<int32_t> & off, std::vector<uint64_t> & mask) {
	const int64_t ns = (nrows + kSliceRows - 1) / kSliceRows;
	const int64_t budget = check_budget ? nnz + nnz / 4 + kSliceRows : INT64_MAX;
	ptr.assign(ns + 1, 0);
	off.clear();
	mask.clear();
	std::vector<int64_t> U;
	int64_t total = 0;
	for (int64_t s = 0; s < ns; ++s) {
		const int64_t r0 = s * kSliceRows, r1 = std::min<int64_t>(nrows, r0 + kSliceRows);
		U.clear();
		for (int64_t r = r0; r < r1; ++r)
			for (I e = rp[r]; e < rp[r + 1]; ++e) {
				if (e > rp[r] && ci[e] <= ci[e - 1]) return false;
				const int64_t d = (int64_t)ci[e] - r;
				if (d < INT32_MIN || d > INT32_MAX) return false;
				U.push_back(d);
			}
		std::sort(U.begin(), U.end());
		U.erase(std::unique(U.begin(), U.end()), U.end());
		if ((int64_t)U.size() > kSliceRows) return false;
		ptr[s] = total;
		total += (int64_t)U.size() * kSliceRows;
		if (total > budget) return false;
		const size_t q0 = off.size();
		for (int64_t d : U) off.push_back((int32_t)d);
		mask.resize(off.size(), 0);
		for (int64_t r = r0; r < r1; ++r) {
			size_t j = 0;
			for (I e = rp[r]; e < rp[r + 1]; ++e) {
				const int64_t d = (int64_t)ci[e] - r;
				while (U[j] != d) ++j;
				mask[q0 + j] |= 1ull << (r - r0);
			}
		}
	}
	ptr[ns] = total;
	return true;
}

template<class I, class V>
static void sdia_fill_t(fsx_ctx * ctx, fsx_csr * A) {
	const int64_t nslots = ((A->nrows + kSliceRows - 1) / kSliceRows) * kSliceRows;
	if (nslots == 0) return;
	const unsigned g = (unsigned)((nslots + 255) / 256);
	hipLaunchKernelGGL((sdia_scatter_kernel<I, V>), dim3(g), dim3(256), 0, ctx->stream,
	                   A->nrows, nslots, (const I *)A->rowptr(), (const I *)A->colind(),
	                   (const V *)A->vals->p, (const int64_t *)A->idx->dg_ptr->p,
	                   (const int32_t *)A->idx->dg_off->p, (V *)A->dg_vals->p);
	FSX_HIP(hipGetLastError());
	ctx->launches++;
}

// Fill A's slice-offset values from its CSR values (stream-ordered).
static void sdia_fill(fsx_ctx * ctx, fsx_csr * A) {
	if (A->it == FSX_I32) {
		if (A->vt == FSX_F64) sdia_fill_t<int32_t, double>(ctx, A);
		else sdia_fill_t<int32_t, float>(ctx, A);
	}
	else {
		if (A->vt == FSX_F64) sdia_fill_t<int64_t, double>(ctx, A);
		else sdia_fill_t<int64_t, float>(ctx, A);
	}
}

bool ensure_sdia(fsx_ctx * ctx, fsx_csr * A, const void * host_rowptr,
                 const void * host_colind, bool check_budget) {
	auto & X = *A->idx;
	if (!X.dg_ptr) {
		if (A->nrows == 0) return false;
		std::vector<int64_t> ptr;
		std::vector<int32_t> off;
		std::vector<uint64_t> mask;
		const bool ok =
			A->it == FSX_I32
				? sdia_analyze<int32_t>(A->nrows, A->nnz, (const int32_t *)host_rowptr,
				                        (const int32_t *)host_colind, check_budget, ptr,
				                        off, mask)
				: sdia_analyze<int64_t>(A->nrows, A->nnz, (const int64_t *)host_rowptr,
				                        (const int64_t *)host_colind, check_budget, ptr,
				                        off, mask);
		if (!ok) return false;
		// padding: the kernels read a full tile of offsets and masks
		off.resize(off.size() + kSliceUnroll, 0);
		mask.resize(mask.size() + kSliceUnroll, 0);
		X.dg_nnz = ptr.back();
		X.dg_ptr = std::make_unique<dbuf>(ptr.size() * sizeof(int64_t));
		X.dg_off = std::make_unique<dbuf>(std::max<size_t>(1, off.size()) * sizeof(int32_t));
		X.dg_mask = std::make_unique<dbuf>(std::max<size_t>(1, mask.size()) * sizeof(uint64_t));
		FSX_HIP(hipMemcpyAsync(X.dg_ptr->p, ptr.data(), ptr.size() * sizeof(int64_t),
		                       hipMemcpyHostToDevice, ctx->stream));
		if (!off.empty()) {
			FSX_HIP(hipMemcpyAsync(X.dg_off->p, off.data(), off.size() * sizeof(int32_t),
			                       hipMemcpyHostToDevice, ctx->stream));
			FSX_HIP(hipMemcpyAsync(X.dg_mask->p, mask.data(), mask.size() * sizeof(uint64_t),
			                       hipMemcpyHostToDevice, ctx->stream));
		}
		FSX_HIP(hipStreamSynchronize(ctx->stream)); // host vectors go out of scope
	}
	if (!A->dg_vals) {
		const size_t vsz = A->vt == FSX_F64 ? 8 : 4;
		A->dg_vals = std::make_unique<dbuf>(std::max<int64_t>(1, X.dg_nnz) * vsz);
		sdia_fill(ctx, A);
		FSX_HIP(hipStreamSynchronize(ctx->stream));
	}
	return true;
}

// ---- symmetric stencil layout ---------------------------------------------

// Host analysis (once per pattern): the ascending union of all rows' column
// offsets, at most kStencilMax and closed under negation; every entry's
// mirror present (row r has column r + o exactly when row r + o has column
// r); the stored slots (offsets >= 0, whole slices) within 25% of the
// entries they hold (+ one slice).  Builds the lane masks per (slice, offset).
template<class I>
static bool stencil_analyze(fsx_ctx * ctx, int64_t nrows, const I * rp, const I * ci,
                            csr_index & X) {
	std::vector<int64_t> offs;
	for (int64_t r = 0; r < nrows; ++r)
		for (I e = rp[r]; e < rp[r + 1]; ++e) {
			// columns strictly ascending: the layout sums in offset order
			if (e > rp[r] && ci[e] <= ci[e - 1]) return false;
			const int64_t o = (int64_t)ci[e] - r;
			if (o < INT32_MIN || o > INT32_MAX) return false; // offsets are stored as int32
			if (std::find(offs.begin(), offs.end(), o) == offs.end()) {
				if ((int)offs.size() == kStencilMax) return false;
				offs.push_back(o);
			}
		}
	if (offs.empty()) return false;
	std::sort(offs.begin(), offs.end());
	const int w = (int)offs.size();
	auto find = [&](int64_t o) {
		return (int)(std::lower_bound(offs.begin(), offs.end(), o) - offs.begin());
	};
	int jm[kStencilMax];
	for (int j = 0; j < w; ++j) {
		jm[j] = find(-offs[j]);
		if (jm[j] == w || offs[jm[j]] != -offs[j]) return false;
	}
	const int nneg = find(0);
	const int wu = w - nneg;
	std::vector<uint8_t> bits((size_t)nrows, 0);
	int64_t upper = 0;
	for (int64_t r = 0; r < nrows; ++r)
		for (I e = rp[r]; e < rp[r + 1]; ++e) {
			const int64_t o = (int64_t)ci[e] - r;
			bits[(size_t)r] |= (uint8_t)(1u << find(o));
			upper += o >= 0;
		}
	for (int64_t r = 0; r < nrows; ++r)
		for (int j = 0; j < w; ++j) {
			if (!((bits[(size_t)r] >> j) & 1)) continue;
			const int64_t rr = r + offs[j];
			if (rr < 0 || rr >= nrows || !((bits[(size_t)rr] >> jm[j]) & 1)) return false;
		}
	const int64_t ns = (nrows + kSliceRows - 1) / kSliceRows;
	if (ns * kSliceRows * wu > upper + upper / 4 + (int64_t)kSliceRows * wu) return false;
	std::vector<uint64_t> mask((size_t)(ns * w + kStencilMax), 0);
	for (int64_t r = 0; r < nrows; ++r)
		for (int j = 0; j < w; ++j)
			if ((bits[(size_t)r] >> j) & 1)
				mask[(size_t)((r / kSliceRows) * w + j)] |= 1ull << (r % kSliceRows);
	X.st_w = w;
	X.st_wu = wu;
	for (int j = 0; j < kStencilMax; ++j) {
		X.st_off[j] = j < w ? (int32_t)offs[j] : 0;
		X.st_ju[j] = j >= w ? 0 : offs[j] >= 0 ? j - nneg : jm[j] - nneg;
	}
	X.st_mask = std::make_unique<dbuf>(mask.size() * sizeof(uint64_t));
	FSX_HIP(hipMemcpyAsync(X.st_mask->p, mask.data(), mask.size() * sizeof(uint64_t),
	                       hipMemcpyHostToDevice, ctx->stream));
	FSX_HIP(hipStreamSynchronize(ctx->stream)); // the host mask goes out of scope
	return true;
}

// Values of the offsets >= 0 of row r (thread r of the padded slice range)
// into the layout: zero the row's slots, then place each entry.  Rows and
// offsets are both ascending, so one forward scan finds them.
template<class I, class V>
__global__ void stencil_fill_kernel(int64_t nrows, int64_t nslots, const I * __restrict__ rp,
                                    const I * __restrict__ ci, const V * __restrict__ v,
                                    stencil_desc sd, V * __restrict__ st_val) {
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= nslots) return;
	const int64_t s = r / kSliceRows;
	const int lane = (int)(r % kSliceRows);
	for (int u = 0; u < sd.wu; ++u) st_val[(s * sd.wu + u) * kSliceRows + lane] = V(0);
	if (r >= nrows) return;
	int j = 0;
	for (I e = rp[r]; e < rp[r + 1]; ++e) {
		const int64_t o = (int64_t)ci[e] - r;
		while (j < sd.w && sd.off[j] != o) ++j;
		if (j == sd.w) return; // not this layout's structure (cannot happen)
		if (o >= 0) st_val[(s * sd.wu + sd.ju[j]) * kSliceRows + lane] = v[e];
		++j;
	}
}

__device__ __forceinline__ uint64_t value_bits(double a) {
	return (uint64_t)__double_as_longlong(a);
}
__device__ __forceinline__ uint64_t value_bits(float a) {
	return (uint64_t)(uint32_t)__float_as_int(a);
}

// bad[0] = 1 when an entry at an offset < 0 differs in any bit from its
// mirror's stored value
template<class I, class V>
__global__ void stencil_sym_check_kernel(int64_t nrows, const I * __restrict__ rp,
                                         const I * __restrict__ ci, const V * __restrict__ v,
                                         stencil_desc sd, const V * __restrict__ st_val,
                                         int * __restrict__ bad) {
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= nrows) return;
	int j = 0;
	for (I e = rp[r]; e < rp[r + 1]; ++e) {
		const int64_t o = (int64_t)ci[e] - r;
		if (o >= 0) break;
		while (j < sd.w && sd.off[j] != o) ++j;
		if (j == sd.w) {
			bad[0] = 1;
			return;
		}
		const int64_t rr = r + o;
		const V m = st_val[((rr / kSliceRows) * sd.wu + sd.ju[j]) * kSliceRows + rr % kSliceRows];
		if (value_bits(m) != value_bits(v[e])) bad[0] = 1;
		++j;
	}
}

// Fill A's stencil values from its CSR values and check their symmetry
// (one host synchronisation); sets A->st_sym.
template<class I, class V>
static void stencil_fill_t(fsx_ctx * ctx, fsx_csr * A) {
	const int64_t nslots = ((A->nrows + kSliceRows - 1) / kSliceRows) * kSliceRows;
	const stencil_desc sd = make_stencil_desc(*A->idx);
	dbuf flag(sizeof(int));
	FSX_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), ctx->stream));
	const unsigned g = (unsigned)((nslots + 255) / 256);
	hipLaunchKernelGGL((stencil_fill_kernel<I, V>), dim3(g), dim3(256), 0, ctx->stream,
	                   A->nrows, nslots, (const I *)A->rowptr(), (const I *)A->colind(),
	                   (const V *)A->vals->p, sd, (V *)A->st_vals->p);
	hipLaunchKernelGGL((stencil_sym_check_kernel<I, V>), dim3(g), dim3(256), 0, ctx->stream,
	                   A->nrows, (const I *)A->rowptr(), (const I *)A->colind(),
	                   (const V *)A->vals->p, sd, (const V *)A->st_vals->p, (int *)flag.p);
	FSX_HIP(hipGetLastError());
	ctx->launches += 2;
	int bad = 0;
	FSX_HIP(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
	FSX_HIP(hipStreamSynchronize(ctx->stream));
	ctx->host_syncs++;
	A->st_sym = bad ? 0 : 1;
}

static void stencil_fill(fsx_ctx * ctx, fsx_csr * A) {
	if (A->it == FSX_I32) {
		if (A->vt == FSX_F64) stencil_fill_t<int32_t, double>(ctx, A);
		else stencil_fill_t<int32_t, float>(ctx, A);
	}
	else {
		if (A->vt == FSX_F64) stencil_fill_t<int64_t, double>(ctx, A);
		else stencil_fill_t<int64_t, float>(ctx, A);
	}
}

// The stencil index (analysed once per pattern, from host copies of the
// structure when given, else downloaded), then A's values in it; true when
// A's values are symmetric and the layout can serve its products.  A's
// stencil values are dropped when they are not.
bool ensure_stencil(fsx_ctx * ctx, fsx_csr * A, const void * host_rowptr,
                    const void * host_colind) {
	csr_index & X = *A->idx;
	if (X.st_state < 0) {
		if (A->nrows == 0 || A->nrows != A->ncols) X.st_state = 0;
		else {
			std::vector<char> hrp, hci;
			if (!host_rowptr || !host_colind) {
				const size_t isz = A->it == FSX_I32 ? 4 : 8;
				hrp.resize((A->nrows + 1) * isz);
				hci.resize(std::max<int64_t>(1, A->nnz) * isz);
				FSX_HIP(hipMemcpyAsync(hrp.data(), A->rowptr(), hrp.size(),
				                       hipMemcpyDeviceToHost, ctx->stream));
				if (A->nnz)
					FSX_HIP(hipMemcpyAsync(hci.data(), A->colind(), A->nnz * isz,
					                       hipMemcpyDeviceToHost, ctx->stream));
				FSX_HIP(hipStreamSynchronize(ctx->stream));
				host_rowptr = hrp.data();
				host_colind = hci.data();
			}
			const bool ok =
				A->it == FSX_I32
					? stencil_analyze<int32_t>(ctx, A->nrows, (const int32_t *)host_rowptr,
					                           (const int32_t *)host_colind, X)
					: stencil_analyze<int64_t>(ctx, A->nrows, (const int64_t *)host_rowptr,
					                           (const int64_t *)host_colind, X);
			X.st_state = ok ? 1 : 0;
		}
	}
	if (X.st_state != 1) return false;
	if (!A->st_vals) {
		const int64_t ns = (A->nrows + kSliceRows - 1) / kSliceRows;
		const size_t vsz = A->vt == FSX_F64 ? 8 : 4;
		A->st_vals = std::make_unique<dbuf>((size_t)(ns * kSliceRows * X.st_wu) * vsz);
		stencil_fill(ctx, A);
	}
	if (A->st_sym != 1) {
		A->st_vals.reset();
		return false;
	}
	return true;
}

void ensure_dg_fresh(fsx_ctx * ctx, const fsx_csr * A) {
	if (!A->dg_stale) return;
	fsx_csr * M = const_cast<fsx_csr *>(A);
	if (M->idx->dg_ptr && M->dg_vals) sdia_fill(ctx, M);
	M->dg_stale = false;
}

void refresh_sliced_values(fsx_ctx * ctx, fsx_csr * A) {
	++A->vals_version;
	if (A->idx->dg_ptr && A->dg_vals) sdia_fill(ctx, A);
	A->dg_stale = false;
	if (A->idx->st_state == 1 && A->st_vals) {
		// refilled in place; values that are no longer symmetric send the
		// products to the slice-offset layout
		stencil_fill(ctx, A);
		if (A->st_sym != 1) {
			A->st_vals.reset();
			if (A->plan_kind == 4) A->plan_kind = 3;
		}
	}
	if (!(A->idx->sl_ptr && A->sl_vals)) return;
	if (A->it == FSX_I32) {
		if (A->vt == FSX_F64) refresh_sliced_t<int32_t, double>(ctx, A);
		else refresh_sliced_t<int32_t, float>(ctx, A);
	}
	else {
		if (A->vt == FSX_F64) refresh_sliced_t<int64_t, double>(ctx, A);
		else refresh_sliced_t<int64_t, float>(ctx, A);
	}
}

// Short rows whose slices share few column offsets (stencils) -> slice
// offsets, or the symmetric stencil layout when all rows draw on one short
// offset list and the values are symmetric; other short rows of similar length -> sliced rows (padding <= 10%,
// addressable with the index type); other short-row matrices -> LDS-staged
// row blocks; long rows -> LANES lanes per row with LANES the next power of
// two of the mean row length (2..64).
void choose_plan(fsx_ctx * ctx, fsx_csr * A, const void * host_rowptr,
                 const void * host_colind) {
	int64_t maxr = 0;
	for (int64_t i = 0; i < A->nrows; ++i) {
		int64_t a, b;
		if (A->it == FSX_I32) {
			a = ((const int32_t *)host_rowptr)[i];
			b = ((const int32_t *)host_rowptr)[i + 1];
		}
		else {
			a = ((const int64_t *)host_rowptr)[i];
			b = ((const int64_t *)host_rowptr)[i + 1];
		}
		maxr = std::max(maxr, b - a);
	}
	A->max_row = maxr;
	A->avg_row = A->nrows ? (double)A->nnz / (double)A->nrows : 0.0;
	const double per_row_budget = (double)kStreamCap / kStreamRows;
	if (A->avg_row <= per_row_budget) {
		if (host_colind && ensure_sdia(ctx, A, host_rowptr, host_colind, true)) {
			// symmetric FP64 stencils: the stencil layout, the slice-offset
			// one kept for a later change of values.  FP32 matrices stay on
			// slice offsets, measured faster there (256^3 FP32 CG K1: 0.119
			// against 0.123-0.130 ms for the stencil variants)
			A->plan_kind =
				A->vt == FSX_F64 && ensure_stencil(ctx, A, host_rowptr, host_colind) ? 4 : 3;
			A->plan_lanes = 0;
			return;
		}
		const int64_t padded =
			A->it == FSX_I32
				? slice_offsets<int32_t>(A->nrows, (const int32_t *)host_rowptr).back()
				: slice_offsets<int64_t>(A->nrows, (const int64_t *)host_rowptr).back();
		const bool fits = A->it == FSX_I64 || padded < ((int64_t)1 << 31);
		if (A->nrows > 0 && fits && padded <= A->nnz + A->nnz / 10 + kSliceRows) {
			A->plan_kind = 2;
			A->plan_lanes = 0;
			ensure_sliced(ctx, A, host_rowptr);
			return;
		}
		A->plan_kind = 0;
		A->plan_lanes = 0;
	}
	else {
		A->plan_kind = 1;
		int l = 2;
		while (l < 64 && l < A->avg_row) l <<= 1;
		A->plan_lanes = l;
	}
}

// x element type follows the vector dtype; ACC follows the matrix values
// (FP64 matrices accumulate in FP64; FP32 matrices with FP32 vectors in FP32,
// with FP64 vectors in FP64).
template<template<class, class> class EpiT, class MakeEpi>
void spmv_dispatch(fsx_ctx * ctx, const fsx_csr * A, const void * x, int xdt,
                   MakeEpi && make) {
	by_matrix_type(A, [&](auto * vtag, auto * itag) {
		using V = std::remove_pointer_t<decltype(vtag)>;
		using I = std::remove_pointer_t<decltype(itag)>;
		if (xdt == FSX_F64) {
			using ACC = double;
			auto epi = make((EpiT<ACC, double> *)nullptr);
			spmv_launch_typed<decltype(epi), ACC, V, I, double>(
				ctx->stream, A, (const double *)x, epi);
		}
		else {
			using ACC = std::conditional_t<std::is_same_v<V, double>, double, float>;
			auto epi = make((EpiT<ACC, float> *)nullptr);
			spmv_launch_typed<decltype(epi), ACC, V, I, float>(
				ctx->stream, A, (const float *)x, epi);
		}
	});
	ctx->launches++;
}

static void check_op(const fsx_csr * A, const fsx_vec * x, const fsx_vec * y) {
	FSX_REQUIRE(A, "null matrix");
	check_vec(x, "x");
	check_vec(y, "y");
	FSX_REQUIRE((int64_t)x->n == A->ncols,
	            "x length " + std::to_string(x->n) + " != ncols " +
	                std::to_string(A->ncols));
	FSX_REQUIRE((int64_t)y->n == A->nrows,
	            "y length " + std::to_string(y->n) + " != nrows " +
	                std::to_string(A->nrows));
	FSX_REQUIRE(x->dt == y->dt, "x and y dtypes differ");
	FSX_REQUIRE(x->ptr() != y->ptr() || A->nrows == 0,
	            "spmv: x and y must be distinct vectors");
}

// ---- transpose structure (for A^T x) ------------------------------------

template<class I>
__global__ void spmv_t_kernel(int64_t ncols, const I * __restrict__ trp,
                              const I * __restrict__ tci,
                              const I * __restrict__ perm,
                              const double * __restrict__ vals,
                              const double * __restrict__ x,
                              double * __restrict__ y) {
	const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= ncols) return;
	double s = 0.0;
	for (I e = trp[j]; e < trp[j + 1]; ++e) s += vals[perm[e]] * x[tci[e]];
	y[j] = s;
}

template<class I>
static void build_transpose(fsx_ctx * ctx, const fsx_csr * A) {
	std::vector<I> rp(A->nrows + 1), ci(A->nnz);
	FSX_HIP(hipMemcpyAsync(rp.data(), A->rowptr(), rp.size() * sizeof(I),
	                       hipMemcpyDeviceToHost, ctx->stream));
	FSX_HIP(hipMemcpyAsync(ci.data(), A->colind(), ci.size() * sizeof(I),
	                       hipMemcpyDeviceToHost, ctx->stream));
	FSX_HIP(hipStreamSynchronize(ctx->stream));
	std::vector<I> trp(A->ncols + 1, 0), tci(A->nnz), perm(A->nnz);
	for (int64_t e = 0; e < A->nnz; ++e) trp[ci[e] + 1]++;
	for (int64_t j = 0; j < A->ncols; ++j) trp[j + 1] += trp[j];
	std::vector<I> fill(trp.begin(), trp.end() - 1);
	for (int64_t i = 0; i < A->nrows; ++i)
		for (I e = rp[i]; e < rp[i + 1]; ++e) {
			I j = ci[e];
			I pos = fill[j]++;
			tci[pos] = (I)i;
			perm[pos] = e;
		}
	auto & X = *A->idx;
	X.t_rowptr = std::make_unique<dbuf>(trp.size() * sizeof(I));
	X.t_colind = std::make_unique<dbuf>(std::max<size_t>(1, tci.size()) * sizeof(I));
	X.t_perm = std::make_unique<dbuf>(std::max<size_t>(1, perm.size()) * sizeof(I));
	FSX_HIP(hipMemcpyAsync(X.t_rowptr->p, trp.data(), trp.size() * sizeof(I),
	                       hipMemcpyHostToDevice, ctx->stream));
	if (A->nnz) {
		FSX_HIP(hipMemcpyAsync(X.t_colind->p, tci.data(), tci.size() * sizeof(I),
		                       hipMemcpyHostToDevice, ctx->stream));
		FSX_HIP(hipMemcpyAsync(X.t_perm->p, perm.data(), perm.size() * sizeof(I),
		                       hipMemcpyHostToDevice, ctx->stream));
	}
	FSX_HIP(hipStreamSynchronize(ctx->stream));
}

// ---- Jacobi relaxation (mg::bound_jacobi::relax, solvers/mg/jacobi.hh:58-93)

template<class I>
__global__ __launch_bounds__(256) void jacobi_relax_kernel(
	int64_t nrows, const I * __restrict__ rp, const I * __restrict__ ci,
	const double * __restrict__ vals, double omega,
	const double * __restrict__ b, const double * __restrict__ tmp,
	double * __restrict__ x) {
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= nrows) return;
	double diag = 0;
	double lpu_x = 0;
	for (I off = rp[r]; off < rp[r + 1]; ++off) {
		const I c = ci[off];
		if ((int64_t)c == r) diag = vals[off];
		else lpu_x += vals[off] * tmp[c];
	}
	const double dinv = 1. / diag;
	x[r] = omega * dinv * (b[r] - lpu_x) + (1 - omega) * tmp[r];
}

template<class I>
__global__ void diag_inv_kernel(int64_t nrows, const I * __restrict__ rp,
                                const I * __restrict__ ci,
                                const double * __restrict__ vals,
                                double * __restrict__ d) {
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= nrows) return;
	double v = 0.0;
	bool found = false;
	for (I off = rp[r]; off < rp[r + 1]; ++off)
		if ((int64_t)ci[off] == r) {
			v = 1.0 / vals[off];
			found = true;
		}
	d[r] = found ? v : 0.0;
}

// a = beta f on every stored entry (coalesced; the diagonal is fixed next)
// a = alpha I + beta f over the CSR values: alpha + beta f at the entries
// whose bit is set in `diag` (each row's diagonal position), beta f
// elsewhere; one coalesced pass
__global__ void shift_scale_csr_kernel(size_t nnz, const uint32_t * __restrict__ diag,
                                       const double * __restrict__ f, double alpha, double beta,
                                       double * __restrict__ a) {
	const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nnz) return;
	const double v = beta * f[e];
	a[e] = ((diag[e >> 5] >> (e & 31)) & 1u) ? alpha + v : v;
}

__global__ void zero_u32_kernel(size_t n, uint32_t * __restrict__ w) {
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) w[i] = 0u;
}

// the bit of each row's diagonal position
__global__ void diag_bits_kernel(int64_t nrows, const int64_t * __restrict__ dpos,
                                 uint32_t * __restrict__ bits) {
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= nrows) return;
	const int64_t e = dpos[r];
	if (e >= 0) atomicOr(&bits[e >> 5], 1u << (e & 31));
}


// The same on the slice-offset layout shared by f and a: slot k = 64 q +
// lane holds row (slice, lane)'s entry at offset off[q] when the lane's mask
// bit is set (absent slots stay zero); the diagonal is offset 0.
__global__ void shift_scale_sdia_kernel(int64_t nslots_total, const uint64_t * __restrict__ mask,
                                        const int32_t * __restrict__ off,
                                        const double * __restrict__ f, double alpha,
                                        double beta, double * __restrict__ a) {
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= nslots_total) return;
	const int64_t q = k / kSliceRows;
	const int lane = (int)(k % kSliceRows);
	if (!((mask[q] >> lane) & 1ull)) return;
	const double v = beta * f[k];
	a[k] = off[q] == 0 ? alpha + v : v;
}

// The same on the stencil layout shared by f and a: slot k holds column
// u = (k / 64) % wu, offset off[w - wu + u], of row (slice k / 64 / wu,
// lane k % 64) when that lane's mask bit is set (absent slots stay zero).
__global__ void shift_scale_stencil_kernel(int64_t nslots_total, const uint64_t * __restrict__ mask,
                                           stencil_desc sd, const double * __restrict__ f,
                                           double alpha, double beta, double * __restrict__ a) {
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= nslots_total) return;
	const int64_t c = k / kSliceRows;
	const int64_t s = c / sd.wu;
	const int j = sd.w - sd.wu + (int)(c % sd.wu);
	const int lane = (int)(k % kSliceRows);
	if (!((mask[s * sd.w + j] >> lane) & 1ull)) return;
	const double v = beta * f[k];
	a[k] = sd.off[j] == 0 ? alpha + v : v;
}

// dpos[r] = position of row r's diagonal entry (the last one if the row
// stores several), -1 if none; missing[0] = 1 when some row has none
template<class I>
__global__ void diag_pos_kernel(int64_t nrows, const I * __restrict__ rp,
                                const I * __restrict__ ci, int64_t * __restrict__ dpos,
                                int * __restrict__ missing) {
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= nrows) return;
	int64_t pos = -1;
	for (I e = rp[r]; e < rp[r + 1]; ++e)
		if ((int64_t)ci[e] == r) pos = (int64_t)e;
	dpos[r] = pos;
	if (pos < 0) missing[0] = 1;
}

// s_r = 1 / sqrt(a_rr) (FP64); bad[0] = 1 when a row has no stored
// diagonal entry or a non-positive one (that row's s is 0)
template<class I>
__global__ void sym_scale_kernel(int64_t nrows, const I * __restrict__ rp,
                                 const I * __restrict__ ci, const double * __restrict__ v,
                                 double * __restrict__ s, int * __restrict__ bad) {
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= nrows) return;
	double d = 0.0;
	for (I e = rp[r]; e < rp[r + 1]; ++e)
		if ((int64_t)ci[e] == r) d = v[e];
	const bool ok = d > 0.0;
	s[r] = ok ? 1.0 / sqrt(d) : 0.0;
	if (!ok && bad) bad[0] = 1;
}

// a32_rc = (float)(a_rc (s_r s_c)): the product of the two scales is the
// same for a_rc and a_cr, so symmetric values give bitwise symmetric ones
template<class I>
__global__ void sym_scaled_f32_kernel(int64_t nrows, const I * __restrict__ rp,
                                      const I * __restrict__ ci,
                                      const double * __restrict__ v,
                                      const double * __restrict__ s, float * __restrict__ a32) {
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= nrows) return;
	const double sr = s[r];
	for (I e = rp[r]; e < rp[r + 1]; ++e) a32[e] = (float)(v[e] * (sr * s[ci[e]]));
}

// missing[0] = 1 when some row has no stored diagonal entry
template<class I>
__global__ void diag_missing_kernel(int64_t nrows, const I * __restrict__ rp,
                                    const I * __restrict__ ci, int * __restrict__ missing) {
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= nrows) return;
	bool found = false;
	for (I e = rp[r]; e < rp[r + 1]; ++e) found |= ((int64_t)ci[e] == r);
	if (!found) missing[0] = 1;
}

template<class S, class D>
__global__ void cast_kernel(size_t n, const S * __restrict__ s, D * __restrict__ d) {
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) d[i] = (D)s[i];
}

__global__ void copy_kernel(size_t n, const double * __restrict__ s,
                            double * __restrict__ d) {
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) d[i] = s[i];
}

void sym_scale_vector(fsx_ctx * ctx, const fsx_csr * A, double * s, int * bad) {
	const int64_t n = A->nrows;
	if (n == 0) return;
	const unsigned g = (unsigned)((n + 255) / 256);
	if (A->it == FSX_I32)
		hipLaunchKernelGGL((sym_scale_kernel<int32_t>), dim3(g), dim3(256), 0, ctx->stream, n,
		                   (const int32_t *)A->rowptr(), (const int32_t *)A->colind(),
		                   (const double *)A->vals->p, s, bad);
	else
		hipLaunchKernelGGL((sym_scale_kernel<int64_t>), dim3(g), dim3(256), 0, ctx->stream, n,
		                   (const int64_t *)A->rowptr(), (const int64_t *)A->colind(),
		                   (const double *)A->vals->p, s, bad);
	FSX_HIP(hipGetLastError());
	ctx->launches++;
}

} // namespace fsx

using namespace fsx;

extern "C" {

int fsx_csr_create(fsx_ctx * ctx, int64_t nrows, int64_t ncols, int64_t nnz,
                   fsx_itype it, const void * rowptr, const void * colind,
                   fsx_dtype vt, const void * vals, fsx_csr ** out) {
	return guard([&] {
		FSX_REQUIRE(ctx, "null context");
		FSX_REQUIRE(out, "null output");
		FSX_REQUIRE(nrows >= 0 && ncols >= 0 && nnz >= 0, "negative size");
		FSX_REQUIRE(it == FSX_I32 || it == FSX_I64, "bad index type");
		FSX_REQUIRE(vt == FSX_F64 || vt == FSX_F32, "bad value type");
		FSX_REQUIRE(rowptr, "null rowptr");
		FSX_REQUIRE(nnz == 0 || (colind && vals), "null colind/vals");
		if (it == FSX_I32)
			FSX_REQUIRE(nnz < ((int64_t)1 << 31) && nrows < ((int64_t)1 << 31) &&
			                ncols < ((int64_t)1 << 31),
			            "int32 indices cannot address this matrix");
		const size_t isz = it == FSX_I32 ? 4 : 8;
		const size_t vsz = vt == FSX_F64 ? 8 : 4;
		// validate structure on the host: monotone rowptr, columns in range.
		auto rp_at = [&](int64_t i) -> int64_t {
			return it == FSX_I32 ? ((const int32_t *)rowptr)[i]
			                     : ((const int64_t *)rowptr)[i];
		};
		FSX_REQUIRE(rp_at(0) == 0, "rowptr[0] must be 0");
		FSX_REQUIRE(rp_at(nrows) == nnz, "rowptr[nrows] must equal nnz");
		for (int64_t i = 0; i < nrows; ++i)
			FSX_REQUIRE(rp_at(i + 1) >= rp_at(i), "rowptr not monotone");
		for (int64_t e = 0; e < nnz; ++e) {
			int64_t c = it == FSX_I32 ? ((const int32_t *)colind)[e]
			                          : ((const int64_t *)colind)[e];
			FSX_REQUIRE(c >= 0 && c < ncols, "column index out of range");
		}
		auto A = std::make_unique<fsx_csr>();
		A->ctx = ctx;
		A->nrows = nrows;
		A->ncols = ncols;
		A->nnz = nnz;
		A->it = it;
		A->vt = vt;
		A->idx = std::make_shared<csr_index>();
		A->idx->rowptr = std::make_unique<dbuf>((nrows + 1) * isz);
		A->idx->colind = std::make_unique<dbuf>(std::max<int64_t>(nnz, 1) * isz);
		A->vals = std::make_unique<dbuf>(std::max<int64_t>(nnz, 1) * vsz);
		FSX_HIP(hipMemcpyAsync(A->idx->rowptr->p, rowptr, (nrows + 1) * isz,
		                       hipMemcpyHostToDevice, ctx->stream));
		if (nnz) {
			FSX_HIP(hipMemcpyAsync(A->idx->colind->p, colind, nnz * isz,
			                       hipMemcpyHostToDevice, ctx->stream));
			FSX_HIP(hipMemcpyAsync(A->vals->p, vals, nnz * vsz,
			                       hipMemcpyHostToDevice, ctx->stream));
		}
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		choose_plan(ctx, A.get(), rowptr, colind);
		ctx_retain(ctx);
		*out = A.release();
	});
}

int fsx_csr_convert(fsx_ctx * ctx, const fsx_csr * A, fsx_dtype vt,
                    fsx_csr ** out) {
	return guard([&] {
		FSX_REQUIRE(ctx && A && out, "null argument");
		ensure_dg_fresh(ctx, A); // B copies A's layouts
		auto B = std::make_unique<fsx_csr>();
		B->ctx = ctx;
		B->nrows = A->nrows;
		B->ncols = A->ncols;
		B->nnz = A->nnz;
		B->it = A->it;
		B->vt = vt;
		B->idx = A->idx; // shared index arrays
		B->plan_kind = A->plan_kind;
		B->plan_lanes = A->plan_lanes;
		B->max_row = A->max_row;
		B->avg_row = A->avg_row;
		const size_t vsz = vt == FSX_F64 ? 8 : 4;
		// m values of src (A's type) into a new buffer of B's type
		auto cast_copy = [&](const dbuf & src, size_t m) {
			auto dst = std::make_unique<dbuf>(std::max<size_t>(m, 1) * vsz);
			if (m == 0) return dst;
			const unsigned g = (unsigned)((m + 255) / 256);
			if (A->vt == FSX_F64 && vt == FSX_F32)
				hipLaunchKernelGGL((cast_kernel<double, float>), dim3(g), dim3(256), 0,
				                   ctx->stream, m, (const double *)src.p, (float *)dst->p);
			else if (A->vt == FSX_F32 && vt == FSX_F64)
				hipLaunchKernelGGL((cast_kernel<float, double>), dim3(g), dim3(256), 0,
				                   ctx->stream, m, (const float *)src.p, (double *)dst->p);
			else
				FSX_HIP(hipMemcpyAsync(dst->p, src.p, m * vsz, hipMemcpyDeviceToDevice,
				                       ctx->stream));
			FSX_HIP(hipGetLastError());
			return dst;
		};
		const size_t avsz = A->vt == FSX_F64 ? 8 : 4;
		B->vals = cast_copy(*A->vals, (size_t)A->nnz);
		// the same sliced, slice-offset and stencil layouts, values converted
		// (padding stays zero); a cast of symmetric values is symmetric
		if (A->sl_vals) B->sl_vals = cast_copy(*A->sl_vals, A->sl_vals->bytes / avsz);
		else if (B->plan_kind == 2)
			B->plan_kind = 0;
		if (A->dg_vals) B->dg_vals = cast_copy(*A->dg_vals, A->dg_vals->bytes / avsz);
		else if (B->plan_kind == 3)
			B->plan_kind = 0;
		// FP32 copies default to slice offsets (as at creation); the stencil
		// values are built on request (fsx_csr_set_plan)
		if (A->st_vals && vt == FSX_F64) {
			B->st_vals = cast_copy(*A->st_vals, A->st_vals->bytes / avsz);
			B->st_sym = A->st_sym;
		}
		else if (B->plan_kind == 4)
			B->plan_kind = B->dg_vals ? 3 : 0;
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		ctx_retain(ctx);
		*out = B.release();
	});
}

int fsx_csr_destroy(fsx_csr * A) {
	return guard([&] {
		if (!A) return;
		fsx_ctx * c = A->ctx;
		(void)hipStreamSynchronize(c->stream);
		// captured solver graphs may hold this matrix's arrays
		invalidate_graphs(c);
		delete A;
		ctx_release(c);
	});
}

int fsx_csr_info(const fsx_csr * A, int64_t * nrows, int64_t * ncols,
                 int64_t * nnz, int * itype, int * vtype) {
	return guard([&] {
		FSX_REQUIRE(A, "null matrix");
		if (nrows) *nrows = A->nrows;
		if (ncols) *ncols = A->ncols;
		if (nnz) *nnz = A->nnz;
		if (itype) *itype = A->it;
		if (vtype) *vtype = A->vt;
	});
}

int fsx_csr_plan(const fsx_csr * A, int * kind, int * lanes) {
	return guard([&] {
		FSX_REQUIRE(A, "null matrix");
		*kind = A->plan_kind;
		*lanes = A->plan_lanes;
	});
}

int fsx_csr_set_plan(fsx_csr * A, int kind, int lanes) {
	return guard([&] {
		FSX_REQUIRE(A, "null matrix");
		FSX_REQUIRE(kind >= 0 && kind <= 4, "plan kind must be 0, 1, 2, 3 or 4");
		if (kind == 1)
			FSX_REQUIRE(lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16 ||
			                lanes == 32 || lanes == 64,
			            "lanes must be a power of two in [2, 64]");
		if (kind == 2) {
			FSX_REQUIRE(A->nrows > 0, "sliced rows need a non-empty matrix");
			ensure_sliced(A->ctx, A, nullptr);
		}
		if (kind == 3) ensure_dg_fresh(A->ctx, A);
		if (kind == 3 && !(A->idx->dg_ptr && A->dg_vals)) {
			FSX_REQUIRE(A->nrows > 0, "slice offsets need a non-empty matrix");
			fsx_ctx * ctx = A->ctx;
			const size_t isz = A->it == FSX_I32 ? 4 : 8;
			std::vector<char> hrp((A->nrows + 1) * isz), hci(std::max<int64_t>(1, A->nnz) * isz);
			FSX_HIP(hipMemcpyAsync(hrp.data(), A->rowptr(), hrp.size(), hipMemcpyDeviceToHost,
			                       ctx->stream));
			if (A->nnz)
				FSX_HIP(hipMemcpyAsync(hci.data(), A->colind(), A->nnz * isz,
				                       hipMemcpyDeviceToHost, ctx->stream));
			FSX_HIP(hipStreamSynchronize(ctx->stream));
			FSX_REQUIRE(ensure_sdia(ctx, A, hrp.data(), hci.data(), false),
			            "matrix structure does not fit the slice-offset layout");
		}
		if (kind == 4) {
			FSX_REQUIRE(A->nrows > 0, "the stencil layout needs a non-empty matrix");
			FSX_REQUIRE(ensure_stencil(A->ctx, A, nullptr, nullptr),
			            A->idx->st_state != 1
			                ? "matrix structure does not fit the symmetric stencil layout"
			                : "matrix values are not symmetric (stencil layout)");
		}
		A->plan_kind = kind;
		A->plan_lanes = kind == 1 ? lanes : 0;
	});
}

int fsx_csr_download(fsx_ctx * ctx, const fsx_csr * A, void * rowptr,
                     void * colind, void * vals) {
	return guard([&] {
		FSX_REQUIRE(ctx && A, "null argument");
		const size_t isz = A->it == FSX_I32 ? 4 : 8;
		const size_t vsz = A->vt == FSX_F64 ? 8 : 4;
		if (rowptr)
			FSX_HIP(hipMemcpyAsync(rowptr, A->rowptr(), (A->nrows + 1) * isz,
			                       hipMemcpyDeviceToHost, ctx->stream));
		if (colind && A->nnz)
			FSX_HIP(hipMemcpyAsync(colind, A->colind(), A->nnz * isz,
			                       hipMemcpyDeviceToHost, ctx->stream));
		if (vals && A->nnz)
			FSX_HIP(hipMemcpyAsync(vals, A->vals->p, A->nnz * vsz,
			                       hipMemcpyDeviceToHost, ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
	});
}

int fsx_csr_spmv(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * x,
                 fsx_vec * y) {
	return guard([&] {
		FSX_REQUIRE(ctx, "null context");
		check_op(A, x, y);
		void * yp = y->ptr();
		if (x->dt == FSX_F64)
			spmv_dispatch<epi_store>(ctx, A, x->ptr(), x->dt, [&](auto * t) {
				using E = std::remove_pointer_t<decltype(t)>;
				E e{};
				e.y = (decltype(e.y))yp;
				return e;
			});
		else
			spmv_dispatch<epi_store>(ctx, A, x->ptr(), x->dt, [&](auto * t) {
				using E = std::remove_pointer_t<decltype(t)>;
				E e{};
				e.y = (decltype(e.y))yp;
				return e;
			});
	});
}

int fsx_csr_residual(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * b,
                     const fsx_vec * x, fsx_vec * r) {
	return guard([&] {
		FSX_REQUIRE(ctx, "null context");
		check_op(A, x, r);
		check_vec(b, "b");
		check_same(b, r);
		FSX_REQUIRE(b->ptr() != r->ptr() || b->n == 0,
		            "residual: b and r must be distinct (use fsx_csr_spmv + "
		            "fsx_subtract for the in-place form)");
		void * rp = r->ptr();
		const void * bp = b->ptr();
		spmv_dispatch<epi_residual>(ctx, A, x->ptr(), x->dt, [&](auto * t) {
			using E = std::remove_pointer_t<decltype(t)>;
			E e{};
			e.r = (decltype(e.r))rp;
			e.b = (decltype(e.b))bp;
			return e;
		});
	});
}

int fsx_spmv_dot(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * p,
                 fsx_vec * w, double * out) {
	return guard([&] {
		FSX_REQUIRE(ctx, "null context");
		check_op(A, p, w);
		FSX_REQUIRE(A->nrows == A->ncols, "spmv_dot needs a square matrix");
		void * wp = w->ptr();
		const void * pp = p->ptr();
		double * slot = ctx->ws.scalars + 1;
		ensure_partials(ctx, spmv_partials_size(A, 1));
		spmv_dispatch<epi_dot>(ctx, A, p->ptr(), p->dt, [&](auto * t) {
			using E = std::remove_pointer_t<decltype(t)>;
			E e{};
			e.w = (decltype(e.w))wp;
			e.p = (decltype(e.p))pp;
			e.partials = ctx->ws.partials;
			return e;
		});
		spmv_finalize<1>(ctx->stream, A, ctx->ws.partials, fin_store<1>{slot});
		FSX_HIP(hipGetLastError());
		*out = fetch_scalar(ctx, slot);
	});
}

int fsx_csr_set_values(fsx_ctx * ctx, fsx_csr * A, const void * vals) {
	return guard([&] {
		FSX_REQUIRE(ctx && A, "null argument");
		FSX_REQUIRE(vals || A->nnz == 0, "null values");
		const size_t vsz = A->vt == FSX_F64 ? 8 : 4;
		if (A->nnz)
			FSX_HIP(hipMemcpyAsync(A->vals->p, vals, (size_t)A->nnz * vsz,
			                       hipMemcpyHostToDevice, ctx->stream));
		refresh_sliced_values(ctx, A);
		FSX_HIP(hipStreamSynchronize(ctx->stream)); // the host buffer may go away
	});
}

int fsx_csr_set_values_f32_sym_scaled(fsx_ctx * ctx, fsx_csr * A32, const fsx_csr * A64) {
	return guard([&] {
		FSX_REQUIRE(ctx && A32 && A64, "null argument");
		FSX_REQUIRE(A32->vt == FSX_F32 && A64->vt == FSX_F64,
		            "set_values_f32_sym_scaled: FP32 destination, FP64 source");
		FSX_REQUIRE(A32->idx == A64->idx,
		            "set_values_f32_sym_scaled: A32 must be fsx_csr_convert(A64, FSX_F32)");
		FSX_REQUIRE(A64->nrows == A64->ncols, "set_values_f32_sym_scaled: square matrix");
		const int64_t n = A64->nrows;
		if (n == 0) return;
		dbuf s((size_t)n * sizeof(double));
		dbuf flag(sizeof(int));
		FSX_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), ctx->stream));
		sym_scale_vector(ctx, A64, (double *)s.p, (int *)flag.p);
		int bad = 0;
		FSX_HIP(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		ctx->host_syncs++;
		FSX_REQUIRE(!bad, "set_values_f32_sym_scaled: a diagonal entry is missing or not positive");
		const unsigned g = (unsigned)((n + 255) / 256);
		if (A64->it == FSX_I32)
			hipLaunchKernelGGL((sym_scaled_f32_kernel<int32_t>), dim3(g), dim3(256), 0,
			                   ctx->stream, n, (const int32_t *)A64->rowptr(),
			                   (const int32_t *)A64->colind(), (const double *)A64->vals->p,
			                   (const double *)s.p, (float *)A32->vals->p);
		else
			hipLaunchKernelGGL((sym_scaled_f32_kernel<int64_t>), dim3(g), dim3(256), 0,
			                   ctx->stream, n, (const int64_t *)A64->rowptr(),
			                   (const int64_t *)A64->colind(), (const double *)A64->vals->p,
			                   (const double *)s.p, (float *)A32->vals->p);
		FSX_HIP(hipGetLastError());
		ctx->launches++;
		refresh_sliced_values(ctx, A32);
		FSX_HIP(hipStreamSynchronize(ctx->stream)); // s and flag are freed on return
	});
}

int fsx_csr_shift_scale(fsx_ctx * ctx, const fsx_csr * F, double alpha, double beta,
                        fsx_csr * A) {
	return guard([&] {
		FSX_REQUIRE(ctx && F && A, "null argument");
		FSX_REQUIRE(A != F, "shift_scale: A must not be F");
		FSX_REQUIRE(A->idx == F->idx,
		            "shift_scale: A must share F's pattern (fsx_csr_convert(F, FSX_F64))");
		FSX_REQUIRE(A->vt == FSX_F64 && F->vt == FSX_F64, "shift_scale: FP64 matrices");
		FSX_REQUIRE(F->nrows == F->ncols, "shift_scale: square matrix");
		const int64_t n = F->nrows;
		if (n == 0) return;
		csr_index & ix = *F->idx;
		const unsigned g = (unsigned)((n + 255) / 256);
		if (!ix.diag_pos) {
			// each row's diagonal position, found once per pattern
			auto pos = std::make_unique<dbuf>((size_t)n * sizeof(int64_t));
			dbuf flag(sizeof(int));
			FSX_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), ctx->stream));
			if (F->it == FSX_I32)
				hipLaunchKernelGGL((diag_pos_kernel<int32_t>), dim3(g), dim3(256), 0,
				                   ctx->stream, n, (const int32_t *)F->rowptr(),
				                   (const int32_t *)F->colind(), (int64_t *)pos->p, (int *)flag.p);
			else
				hipLaunchKernelGGL((diag_pos_kernel<int64_t>), dim3(g), dim3(256), 0,
				                   ctx->stream, n, (const int64_t *)F->rowptr(),
				                   (const int64_t *)F->colind(), (int64_t *)pos->p, (int *)flag.p);
			FSX_HIP(hipGetLastError());
			int missing = 0;
			FSX_HIP(hipMemcpyAsync(&missing, flag.p, sizeof(int), hipMemcpyDeviceToHost,
			                       ctx->stream));
			FSX_HIP(hipStreamSynchronize(ctx->stream));
			ctx->host_syncs++;
			ix.diag_full = missing ? 0 : 1;
			ix.diag_pos = std::move(pos);
			// the diagonal positions as one bit per entry
			const size_t words = (size_t)F->nnz / 32 + 1;
			auto bits = std::make_unique<dbuf>(words * sizeof(uint32_t));
			hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)((words + 255) / 256)), dim3(256), 0,
			                   ctx->stream, words, (uint32_t *)bits->p);
			hipLaunchKernelGGL(diag_bits_kernel, dim3(g), dim3(256), 0, ctx->stream, n,
			                   (const int64_t *)ix.diag_pos->p, (uint32_t *)bits->p);
			FSX_HIP(hipGetLastError());
			ix.diag_bits = std::move(bits);
		}
		FSX_REQUIRE(ix.diag_full == 1 || alpha == 0.0,
		            "shift_scale: a row of F has no stored diagonal entry");
		const size_t nnz = (size_t)F->nnz;
		if (nnz) {
			hipLaunchKernelGGL(shift_scale_csr_kernel, dim3((unsigned)((nnz + 255) / 256)),
			                   dim3(256), 0, ctx->stream, nnz, (const uint32_t *)ix.diag_bits->p,
			                   (const double *)F->vals->p, alpha, beta, (double *)A->vals->p);
			FSX_HIP(hipGetLastError());
			ctx->launches += 1;
		}
		// the SpMV layouts: the stencil one and the slice-offset one rewritten
		// in place from F's, the sliced one re-copied from the CSR values.  A
		// matrix whose products run on the stencil layout (and stay there:
		// F's stencil values exist, so they are symmetric) leaves its
		// slice-offset values behind, rebuilt when next needed.
		++A->vals_version;
		const bool st_serves = A->plan_kind == 4 && ix.st_state == 1 && A->st_vals && F->st_vals;
		if (ix.dg_ptr && A->dg_vals) {
			if (st_serves)
				A->dg_stale = true;
			else if (F->dg_vals && ix.dg_nnz > 0) {
				ensure_dg_fresh(ctx, F);
				A->dg_stale = false;
				const int64_t t = ix.dg_nnz;
				hipLaunchKernelGGL(shift_scale_sdia_kernel, dim3((unsigned)((t + 255) / 256)),
				                   dim3(256), 0, ctx->stream, t, (const uint64_t *)ix.dg_mask->p,
				                   (const int32_t *)ix.dg_off->p, (const double *)F->dg_vals->p,
				                   alpha, beta, (double *)A->dg_vals->p);
				FSX_HIP(hipGetLastError());
				ctx->launches++;
			}
			else {
				sdia_fill(ctx, A);
				A->dg_stale = false;
			}
		}
		if (ix.st_state == 1 && A->st_vals) {
			if (F->st_vals) {
				// F's stencil values exist only when they are symmetric, and
				// alpha I + beta F keeps that
				const int64_t t = (int64_t)(A->st_vals->bytes / sizeof(double));
				hipLaunchKernelGGL(shift_scale_stencil_kernel, dim3((unsigned)((t + 255) / 256)),
				                   dim3(256), 0, ctx->stream, t, (const uint64_t *)ix.st_mask->p,
				                   make_stencil_desc(ix), (const double *)F->st_vals->p, alpha,
				                   beta, (double *)A->st_vals->p);
				FSX_HIP(hipGetLastError());
				ctx->launches++;
				A->st_sym = 1;
			}
			else
				stencil_fill(ctx, A);
			if (A->st_sym != 1) {
				A->st_vals.reset();
				if (A->plan_kind == 4) A->plan_kind = 3;
			}
		}
		if (ix.sl_ptr && A->sl_vals) {
			if (A->it == FSX_I32) refresh_sliced_t<int32_t, double>(ctx, A);
			else refresh_sliced_t<int64_t, double>(ctx, A);
		}
	});
}

int fsx_csr_set_values_f32_from_f64(fsx_ctx * ctx, fsx_csr * A32, const fsx_csr * A64) {
	return guard([&] {
		FSX_REQUIRE(ctx && A32 && A64, "null argument");
		FSX_REQUIRE(A32->vt == FSX_F32 && A64->vt == FSX_F64,
		            "set_values_f32_from_f64: FP32 destination, FP64 source");
		FSX_REQUIRE(A32->idx == A64->idx,
		            "set_values_f32_from_f64: A32 must be fsx_csr_convert(A64, FSX_F32)");
		const size_t n = (size_t)A64->nnz;
		if (n) {
			const unsigned g = (unsigned)((n + 255) / 256);
			hipLaunchKernelGGL((cast_kernel<double, float>), dim3(g), dim3(256), 0,
			                   ctx->stream, n, (const double *)A64->vals->p,
			                   (float *)A32->vals->p);
			FSX_HIP(hipGetLastError());
			ctx->launches++;
		}
		refresh_sliced_values(ctx, A32);
	});
}

int fsx_csr_diag_inv(fsx_ctx * ctx, const fsx_csr * A, fsx_vec * dinv) {
	return guard([&] {
		FSX_REQUIRE(ctx && A, "null argument");
		check_vec(dinv, "dinv");
		FSX_REQUIRE((int64_t)dinv->n == A->nrows, "dinv length != nrows");
		FSX_REQUIRE(A->vt == FSX_F64, "diag_inv needs an FP64 matrix");
		if (A->nrows == 0) return;
		unsigned g = (unsigned)((A->nrows + 255) / 256);
		if (dinv->dt == FSX_F64) {
			if (A->it == FSX_I32)
				hipLaunchKernelGGL((diag_inv_kernel<int32_t>), dim3(g), dim3(256), 0,
				                   ctx->stream, A->nrows, (const int32_t *)A->rowptr(),
				                   (const int32_t *)A->colind(),
				                   (const double *)A->vals->p, (double *)dinv->ptr());
			else
				hipLaunchKernelGGL((diag_inv_kernel<int64_t>), dim3(g), dim3(256), 0,
				                   ctx->stream, A->nrows, (const int64_t *)A->rowptr(),
				                   (const int64_t *)A->colind(),
				                   (const double *)A->vals->p, (double *)dinv->ptr());
		}
		else {
			// FP32 preconditioner vector: 1/a_ii formed in FP64, then rounded
			dbuf tmp(A->nrows * 8);
			if (A->it == FSX_I32)
				hipLaunchKernelGGL((diag_inv_kernel<int32_t>), dim3(g), dim3(256), 0,
				                   ctx->stream, A->nrows, (const int32_t *)A->rowptr(),
				                   (const int32_t *)A->colind(),
				                   (const double *)A->vals->p, (double *)tmp.p);
			else
				hipLaunchKernelGGL((diag_inv_kernel<int64_t>), dim3(g), dim3(256), 0,
				                   ctx->stream, A->nrows, (const int64_t *)A->rowptr(),
				                   (const int64_t *)A->colind(),
				                   (const double *)A->vals->p, (double *)tmp.p);
			hipLaunchKernelGGL((cast_kernel<double, float>), dim3(g), dim3(256), 0,
			                   ctx->stream, (size_t)A->nrows, (const double *)tmp.p,
			                   (float *)dinv->ptr());
			FSX_HIP(hipStreamSynchronize(ctx->stream));
		}
		FSX_HIP(hipGetLastError());
		ctx->launches++;
	});
}

int fsx_jacobi_relax(fsx_ctx * ctx, const fsx_csr * A, float omega, int nrelax,
                     const fsx_vec * b, fsx_vec * x) {
	return guard([&] {
		FSX_REQUIRE(ctx && A, "null argument");
		check_vec(b, "b");
		check_vec(x, "x");
		FSX_REQUIRE(A->vt == FSX_F64 && x->dt == FSX_F64 && b->dt == FSX_F64,
		            "jacobi_relax is FP64");
		FSX_REQUIRE(A->nrows == A->ncols && (int64_t)x->n == A->nrows &&
		                (int64_t)b->n == A->nrows,
		            "jacobi_relax: shape mismatch");
		FSX_REQUIRE(nrelax >= 0, "nrelax must be >= 0");
		FSX_REQUIRE(b->ptr() != x->ptr() || A->nrows == 0, "jacobi_relax: b and x must differ");
		if (A->nrows == 0 || nrelax == 0) return;
		const size_t n = (size_t)A->nrows;
		double * tmp = ensure_relax_buffer(ctx, n);
		const unsigned g = (unsigned)((n + 255) / 256);
		double * xp = (double *)x->ptr();
		if (A->plan_kind == 2 || A->plan_kind == 3 || A->plan_kind == 4) {
			// sweeps alternate between x and the buffer and end in x: an odd
			// count starts from a copy of x (jacobi.hh:63's tmp = x)
			double * src = xp;
			double * dst = tmp;
			if (nrelax % 2) {
				hipLaunchKernelGGL(copy_kernel, dim3(g), dim3(256), 0, ctx->stream, n,
				                   (const double *)xp, tmp);
				std::swap(src, dst);
			}
			for (int s = 0; s < nrelax; ++s) {
				epi_relax e{(double)omega, (const double *)b->ptr(), dst};
				by_matrix_type(A, [&](auto * vtag, auto * itag) {
					using V = std::remove_pointer_t<decltype(vtag)>;
					using I = std::remove_pointer_t<decltype(itag)>;
					if constexpr (std::is_same_v<V, double>)
						spmv_launch_typed<epi_relax, double, V, I, double>(ctx->stream, A,
						                                                    (const double *)src, e);
				});
				std::swap(src, dst);
			}
			FSX_HIP(hipGetLastError());
			ctx->launches += nrelax + (nrelax % 2);
			return;
		}
		for (int s = 0; s < nrelax; ++s) {
			// tmp = x (jacobi.hh:63), then one sweep
			hipLaunchKernelGGL(copy_kernel, dim3(g), dim3(256), 0, ctx->stream, n,
			                   (const double *)xp, tmp);
			if (A->it == FSX_I32)
				hipLaunchKernelGGL((jacobi_relax_kernel<int32_t>), dim3(g), dim3(256),
				                   0, ctx->stream, A->nrows,
				                   (const int32_t *)A->rowptr(),
				                   (const int32_t *)A->colind(),
				                   (const double *)A->vals->p, (double)omega,
				                   (const double *)b->ptr(), (const double *)tmp, xp);
			else
				hipLaunchKernelGGL((jacobi_relax_kernel<int64_t>), dim3(g), dim3(256),
				                   0, ctx->stream, A->nrows,
				                   (const int64_t *)A->rowptr(),
				                   (const int64_t *)A->colind(),
				                   (const double *)A->vals->p, (double)omega,
				                   (const double *)b->ptr(), (const double *)tmp, xp);
			FSX_HIP(hipGetLastError());
			ctx->launches += 2;
		}
	});
}

int fsx_csr_spmv_t(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * x,
                   fsx_vec * y) {
	return guard([&] {
		FSX_REQUIRE(ctx && A, "null argument");
		check_vec(x, "x");
		check_vec(y, "y");
		FSX_REQUIRE(A->vt == FSX_F64 && x->dt == FSX_F64 && y->dt == FSX_F64,
		            "spmv_t is FP64");
		FSX_REQUIRE((int64_t)x->n == A->nrows && (int64_t)y->n == A->ncols,
		            "spmv_t: shape mismatch");
		FSX_REQUIRE(x->ptr() != y->ptr(), "spmv_t: x and y must differ");
		if (!A->idx->t_rowptr) {
			if (A->it == FSX_I32) build_transpose<int32_t>(ctx, A);
			else build_transpose<int64_t>(ctx, A);
		}
		if (A->ncols == 0) return;
		unsigned g = (unsigned)((A->ncols + 255) / 256);
		auto & X = *A->idx;
		if (A->it == FSX_I32)
			hipLaunchKernelGGL((spmv_t_kernel<int32_t>), dim3(g), dim3(256), 0,
			                   ctx->stream, A->ncols, (const int32_t *)X.t_rowptr->p,
			                   (const int32_t *)X.t_colind->p,
			                   (const int32_t *)X.t_perm->p,
			                   (const double *)A->vals->p, (const double *)x->ptr(),
			                   (double *)y->ptr());
		else
			hipLaunchKernelGGL((spmv_t_kernel<int64_t>), dim3(g), dim3(256), 0,
			                   ctx->stream, A->ncols, (const int64_t *)X.t_rowptr->p,
			                   (const int64_t *)X.t_colind->p,
			                   (const int64_t *)X.t_perm->p,
			                   (const double *)A->vals->p, (const double *)x->ptr(),
			                   (double *)y->ptr());
		FSX_HIP(hipGetLastError());
		ctx->launches++;
	});
}

// ---- timing --------------------------------------------------------------

int fsx_csr_layout_bytes(const fsx_csr * A, size_t x_bytes, uint64_t * bytes) {
	return guard([&] {
		FSX_REQUIRE(A && bytes, "null argument");
		const uint64_t vb = A->vt == FSX_F64 ? 8 : 4, ib = A->it == FSX_I32 ? 4 : 8;
		const uint64_t n = (uint64_t)A->nrows, nsl = (n + kSliceRows - 1) / kSliceRows;
		const uint64_t vec = n * x_bytes * 2; // x read once, y written once
		const csr_index & X = *A->idx;
		if (A->plan_kind == 4) {
			// value slots of the offsets >= 0; per slice one 8-byte lane mask
			// per offset (entries at offsets < 0 re-read stored slots)
			*bytes = nsl * kSliceRows * (uint64_t)X.st_wu * vb + nsl * (uint64_t)X.st_w * 8 + vec;
		}
		else if (A->plan_kind == 3) {
			// value slots; per slice one 8-byte pointer; per offset entry (64
			// slots) a 4-byte offset and an 8-byte lane mask
			const uint64_t slots = (uint64_t)X.dg_nnz;
			*bytes = slots * vb + (nsl + 1) * 8 + slots / kSliceRows * 12 + vec;
		}
		else if (A->plan_kind == 2) {
			const uint64_t slots = (uint64_t)X.sl_nnz;
			*bytes = slots * (vb + ib) + (nsl + 1) * ib + vec;
		}
		else
			*bytes = (uint64_t)A->nnz * (vb + ib) + (n + 1) * ib + vec;
	});
}

int fsx_time_spmv(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * x,
                  fsx_vec * y, int reps, double * avg_ms) {
	return guard([&] {
		FSX_REQUIRE(ctx && avg_ms, "null argument");
		check_op(A, x, y);
		FSX_REQUIRE(reps > 0, "reps must be > 0");
		hipEvent_t e0, e1;
		FSX_HIP(hipEventCreate(&e0));
		FSX_HIP(hipEventCreate(&e1));
		FSX_HIP(hipEventRecord(e0, ctx->stream));
		for (int i = 0; i < reps; ++i) {
			int rc = fsx_csr_spmv(ctx, A, x, y);
			if (rc) throw error(rc, get_error());
		}
		FSX_HIP(hipEventRecord(e1, ctx->stream));
		FSX_HIP(hipEventSynchronize(e1));
		float ms = 0;
		FSX_HIP(hipEventElapsedTime(&ms, e0, e1));
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		*avg_ms = (double)ms / reps;
	});
}

int fsx_time_copy(fsx_ctx * ctx, const fsx_vec * src, fsx_vec * dst, int reps,
                  double * avg_ms) {
	return guard([&] {
		FSX_REQUIRE(ctx && avg_ms, "null argument");
		check_vec(src, "src");
		check_vec(dst, "dst");
		check_same(src, dst);
		FSX_REQUIRE(src->dt == FSX_F64, "copy timing uses FP64 vectors");
		FSX_REQUIRE(reps > 0, "reps must be > 0");
		size_t n = src->n;
		unsigned g = (unsigned)((n + 255) / 256);
		hipEvent_t e0, e1;
		FSX_HIP(hipEventCreate(&e0));
		FSX_HIP(hipEventCreate(&e1));
		FSX_HIP(hipEventRecord(e0, ctx->stream));
		for (int i = 0; i < reps; ++i)
			hipLaunchKernelGGL(copy_kernel, dim3(g), dim3(256), 0, ctx->stream, n,
			                   (const double *)src->ptr(), (double *)dst->ptr());
		FSX_HIP(hipGetLastError());
		FSX_HIP(hipEventRecord(e1, ctx->stream));
		FSX_HIP(hipEventSynchronize(e1));
		float ms = 0;
		FSX_HIP(hipEventElapsedTime(&ms, e0, e1));
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		*avg_ms = (double)ms / reps;
	});
}

} // extern "C"
