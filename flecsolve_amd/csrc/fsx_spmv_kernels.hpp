// fsx_spmv_kernels.hpp -- CSR SpMV kernel templates for gfx950.
//
// Reference operation: compressed_ops::spmv (flecsolve/matrices/seq.hh:177-193)
//     y[i] = 0; for off in [rowptr[i], rowptr[i+1]): y[i] += v[off]*x[col[off]]
// reached through mat::parcsr's apply (matrices/parcsr.hh:60-91), which at one
// colour is this single local product.
//
// Two row mappings, chosen per matrix from its row-length profile:
//
//  * Row blocks staged through LDS (short rows; the 5/7-point stencils).
//    A 256-thread workgroup owns 256 consecutive rows.  Its products
//    v[off]*x[col[off]] are formed with one lane per nonzero -- so the value
//    and column streams are read fully coalesced -- and parked in LDS; each
//    thread then sums its own row's products from LDS in CSR order.  This is
//    the CSR-Stream scheme of Greathouse & Daga (SC'14).  Because every row is
//    summed serially from 0 in CSR order with separate multiply and add
//    roundings, y is bit-identical to the reference's serial loop.  A block
//    whose nonzeros overflow the LDS budget falls back to one thread per row
//    (serial, uncoalesced) for that block only, so the order is kept there
//    too.
//
//  * Sub-wavefront vector per row (long rows): LANES lanes (2..64) share a
//    row, stride through it, and combine with cross-lane shuffles.  The
//    summation order differs from the serial loop (results agree to rounding).
//
//  * Sliced rows (short rows of similar length; the default for stencils):
//    the CSR is re-laid out once on the device in slices of 64 rows (one
//    wavefront), entry-major within a slice, each slice padded to its
//    longest row (sliced ELLPACK with slice height = wavefront width).  Lane
//    l owns row 64 s + l and reads its j-th entry at sl_ptr[s] + 64 j + l, so
//    every value/column load of a wavefront is one contiguous 512/256-byte
//    segment with no LDS round trip or barrier.  Entries keep their CSR
//    order and each row is summed serially, so y is bit-identical to the
//    reference's loop.  Measured on the 256^3 7-point matrix this moves the
//    kernel from ~64% (best LDS-staged CSR variant) to ~70% of HBM peak.
//
//  * Symmetric stencil (structurally and bitwise symmetric matrices whose
//    rows all draw on one list of at most 8 column offsets: the 3-D 7-point
//    and 2-D 5-point stencils): the offsets travel as kernel arguments, each
//    64-row slice keeps one lane mask per offset, and values are stored only
//    for the offsets >= 0 (the diagonal and the entries above it); the entry
//    of row r at a negative offset o is read from row r + o's stored entry at
//    -o.  FP64 at 256^3: y = A x 0.227 -> 0.198 ms and the CG K1 0.232 ->
//    0.200 ms against the slice-offset layout; the FP32 CG K1 measured
//    0.123-0.184 ms against 0.119-0.124 there across launch shapes, so FP32
//    matrices stay on slice offsets.  Measured and not kept: mirroring only
//    the offsets that are whole slices (-256, -65536; the -1 entry stored,
//    so no mirror load straddles lanes): FP32 0.159 ms, FP64 0.206 ms;
//    non-temporal loads for the last use of a value (mirror reads, the
//    diagonal): FP32 0.157 ms, FP64 K1 0.214 ms; XCD-major block order:
//    FP64 0.202 / 0.209 ms; other shapes for the FP64 K1 (256- or
//    1024-thread blocks one slice per wave, 256 or 512 threads striding over
//    2-4 slices): 0.196-0.207 ms against 0.196 ms, steady state, for a plain
//    product of 0.192 ms; x at -1 and +1 and the -1 mirror value taken from
//    the neighbouring lanes (DPP wave shifts, one uniform load per slice
//    edge; three of the fourteen loads a row issues): product 0.197 ->
//    0.229 ms and CG K1 0.200 -> 0.513 ms (the reducing instance at 8 waves
//    per SIMD).  Rows keep their CSR order: bit-identical to the serial
//    loop.
//
// Workgroups are dealt to XCDs round-robin by block id.  Measured: keeping
// that default order is faster than remapping blocks to give each XCD a
// contiguous row range (all XCDs then stream neighbouring rows at once), and
// no faster for the slice-offset kernel than grouping runs of 4-32
// consecutive blocks on one XCD (so slices +-n rows apart share an L2).  For
// the slice-offset kernel at 256^3, numbering blocks XCD-major with one
// contiguous run of slices each (every XCD sweeping one eighth of the rows)
// gave the FP32 CG product 0.122 -> 0.119 ms but the FP64 product 0.228 ->
// 0.245 ms and the FP64 CG product 0.232 -> 0.252 ms; not used.  Skipping
// the value loads of a unit diagonal (the scaled FP32 matrix) through a
// wave-uniform address select cost more issue time than the 8% of bytes it
// saved (FP32 CG product 0.120 -> 0.127 ms, and the FP64 products slower
// too); a branch per load split the merged scalar loads of offsets and
// masks; not used.  Taking x at the offsets -1 and +1 from the neighbouring
// lanes of the aligned x[r] load (lane shifts, one broadcast load per slice
// edge) instead of gathers that straddle a cache line: FP32 CG product
// 0.124 -> 0.143 ms with DPP wave shifts, 0.148 ms with ds_bpermute; not
// used.  Issue order within a tile (same loads, same sums): the gathers
// before the value loads 0.120 -> 0.125 ms, the -1/+1 gathers after the
// value loads (so the x[r] lines they share are no longer pending)
// 0.137 ms; values first, then the gathers in offset order, is kept.  Taking a wave's slices two at a time with every load of
// both issued before the first gather (more loads in flight per wave)
// raised the FP32 CG product's registers to 83 (0.141 ms) or, held to 64,
// gave 0.125 ms against 0.118; not used.
//
// The kernels take an epilogue policy `Epi` that consumes each row's sum
// (row() *adds* its dot-product terms to d: a thread may own several rows):
// plain store, residual b - Ax, or store plus fused dot-product terms whose
// per-block partials a one-block finalize kernel reduces next (device-side
// scalar updates of the fused CG happen there).  Epi::skip() lets captured
// iterations become no-ops after the device-side convergence flag is set.
#pragma once

#include "fsx_internal.hpp"

#include <algorithm>
#include <type_traits>

namespace fsx {

// Epilogues with `offdiag = true` receive the row sum without the diagonal
// entry (the remaining entries still in CSR order) and the diagonal value:
// row_offdiag(i, sum, diag, x_i).  Only the sliced and slice-offset kernels
// serve them.
template<class E, class = void>
struct is_offdiag : std::false_type {};
template<class E>
struct is_offdiag<E, std::void_t<decltype(E::offdiag)>> : std::bool_constant<E::offdiag> {};

constexpr int kSpmvThreads = 256;
constexpr int kStreamRows = 256;   // rows per workgroup (LDS-staged mapping)
constexpr int kStreamCap = 2048;   // nonzeros staged per workgroup

template<class T>
__device__ __forceinline__ T ldnt(const T * p) {
	return __builtin_nontemporal_load(p);
}

// Element c of p through a 32-bit byte offset from the (uniform) base: one
// address register per lane instead of two.  Valid while c * sizeof(T)
// stays below 2^32.
template<class T>
__device__ __forceinline__ T ld32(const T * p, uint32_t c) {
	return *(const T *)((const char *)p + c * (uint32_t)sizeof(T));
}

// Gather policies: the value of the multiplied vector at column c (at32:
// the same through a 32-bit offset).
template<class X>
struct gather_x {
	using value_type = X;
	const X * x;
	__device__ __forceinline__ void init() {}
	__device__ __forceinline__ X operator()(int64_t c) const { return x[c]; }
	__device__ __forceinline__ X at32(uint32_t c) const { return ld32(x, c); }
};

// Sum of the thread's row for the LDS-staged mapping.  Valid when
// threadIdx.x < nr.  `srow` holds nr+1 row pointers, `prod` kStreamCap slots.
template<class ACC, class V, class I, class G>
__device__ __forceinline__ ACC stream_rows(int64_t r0, int nr,
                                           const I * __restrict__ rowptr,
                                           const I * __restrict__ colind,
                                           const V * __restrict__ vals,
                                           const G & x, I * srow,
                                           ACC * prod) {
	const int tid = threadIdx.x;
	for (int t = tid; t <= nr; t += kSpmvThreads) srow[t] = rowptr[r0 + t];
	__syncthreads();
	const I base = srow[0];
	const int64_t cnt = (int64_t)(srow[nr] - base);
	ACC sum = ACC(0);
	if (cnt <= kStreamCap) {
		constexpr int NJ = kStreamCap / kSpmvThreads;
		I cols[NJ];
		V vv[NJ];
#pragma unroll
		for (int j = 0; j < NJ; ++j) {
			const int k = tid + j * kSpmvThreads;
			if (k < cnt) {
				cols[j] = ldnt(colind + base + k);
				vv[j] = ldnt(vals + base + k);
			}
		}
#pragma unroll
		for (int j = 0; j < NJ; ++j) {
			const int k = tid + j * kSpmvThreads;
			if (k < cnt) prod[k] = (ACC)vv[j] * (ACC)x(cols[j]);
		}
		__syncthreads();
		if (tid < nr) {
			const int lo = (int)(srow[tid] - base), hi = (int)(srow[tid + 1] - base);
			ACC s = ACC(0);
			for (int e = lo; e < hi; ++e) s += prod[e];
			sum = s;
		}
	}
	else if (tid < nr) {
		// overflow (the block's entries do not fit the stage): each thread
		// sums its own row serially, keeping the CSR order
		const I lo = srow[tid], hi = srow[tid + 1];
		ACC s = ACC(0);
		for (I e = lo; e < hi; ++e) s += (ACC)ldnt(vals + e) * (ACC)x(ldnt(colind + e));
		sum = s;
	}
	return sum;
}

// LANES lanes per row; returns the row sum in every lane of the group.
template<int LANES, class ACC, class V, class I, class G>
__device__ __forceinline__ ACC vector_row(int64_t row, int64_t nrows,
                                          const I * __restrict__ rowptr,
                                          const I * __restrict__ colind,
                                          const V * __restrict__ vals,
                                          const G & x) {
	const int sub = threadIdx.x & (LANES - 1);
	ACC s = ACC(0);
	if (row < nrows) {
		const I lo = rowptr[row], hi = rowptr[row + 1];
		for (I e = lo + sub; e < hi; e += LANES)
			s += (ACC)ldnt(vals + e) * (ACC)x(ldnt(colind + e));
	}
#pragma unroll
	for (int o = LANES / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LANES);
	return s;
}

// ---- kernels -------------------------------------------------------------

template<class Epi, class ACC, class V, class I, class G>
__global__ __launch_bounds__(kSpmvThreads) void spmv_stream_kernel(
	int64_t nrows, const I * __restrict__ rowptr, const I * __restrict__ colind,
	const V * __restrict__ vals, G x, Epi epi) {
	if (epi.skip()) return;
	x.init();
	__shared__ I srow[kStreamRows + 1];
	__shared__ ACC prod[kStreamCap];
	const unsigned blk = blockIdx.x;
	const int64_t r0 = (int64_t)blk * kStreamRows;
	const int nr = (int)min((int64_t)kStreamRows, nrows - r0);
	ACC sum = stream_rows<ACC>(r0, nr, rowptr, colind, vals, x, srow, prod);
	double d[Epi::NV];
#pragma unroll
	for (int i = 0; i < Epi::NV; ++i) d[i] = 0.0;
	if ((int)threadIdx.x < nr) {
		const int64_t i = r0 + threadIdx.x;
		epi.row(i, sum, d, [&] { return x(i); });
	}
	epi.block_end(d, blk, gridDim.x);
}

template<int LANES, class Epi, class ACC, class V, class I, class G>
__global__ __launch_bounds__(kSpmvThreads) void spmv_vector_kernel(
	int64_t nrows, const I * __restrict__ rowptr, const I * __restrict__ colind,
	const V * __restrict__ vals, G x, Epi epi) {
	if (epi.skip()) return;
	x.init();
	constexpr int RPB = kSpmvThreads / LANES;
	const unsigned blk = blockIdx.x;
	const int64_t row = (int64_t)blk * RPB + threadIdx.x / LANES;
	ACC sum = vector_row<LANES, ACC>(row, nrows, rowptr, colind, vals, x);
	double d[Epi::NV];
#pragma unroll
	for (int i = 0; i < Epi::NV; ++i) d[i] = 0.0;
	if ((threadIdx.x & (LANES - 1)) == 0 && row < nrows)
		epi.row(row, sum, d, [&] { return x(row); });
	epi.block_end(d, blk, gridDim.x);
}

constexpr int kSliceRows = 64;   // slice height = wavefront width
constexpr int kSliceUnroll = 8;  // entries loaded ahead of the gathers

// Grid-stride over rows: at most kSlicedMaxBlocks workgroups, so a fused
// dot product leaves few per-block partials for its finalize kernel.
constexpr unsigned kSlicedMaxBlocks = 8192;

template<class Epi, class ACC, class V, class I, class G>
__global__ __launch_bounds__(kSpmvThreads) void spmv_sliced_kernel(
	int64_t nrows, const I * __restrict__ sl_ptr, const I * __restrict__ sl_col,
	const V * __restrict__ sl_val, G x, Epi epi) {
	if (epi.skip()) return;
	x.init();
	const int lane = threadIdx.x & (kSliceRows - 1);
	double d[Epi::NV];
#pragma unroll
	for (int i = 0; i < Epi::NV; ++i) d[i] = 0.0;
	const int64_t stride = (int64_t)gridDim.x * kSpmvThreads;
	for (int64_t r = (int64_t)blockIdx.x * kSpmvThreads + threadIdx.x; r < nrows;
	     r += stride) {
		const int64_t s = r / kSliceRows;
		const I base = sl_ptr[s];
		const int w = (int)((sl_ptr[s + 1] - base) / kSliceRows);
		const I * cp = sl_col + base + lane;
		const V * vp = sl_val + base + lane;
		ACC sum = ACC(0);
		// x at the row's own column, captured from the diagonal entry's gather
		// (epilogues that need it then do not read x again)
		typename G::value_type xself{};
		bool has_self = false;
		double dgv = 0.; // diagonal value (offdiag epilogues)
		// Branch-free body as in the slice-offset kernel: all loads of a tile
		// first, padding (column -1) and slots past the width excluded by a
		// select.
		for (int j0 = 0; j0 < w; j0 += kSliceUnroll) {
			I c[kSliceUnroll];
			V v[kSliceUnroll];
			typename G::value_type xv[kSliceUnroll];
#pragma unroll
			for (int k = 0; k < kSliceUnroll; ++k) {
				const int j = min(j0 + k, w - 1);
				c[k] = ldnt(cp + j * kSliceRows);
				v[k] = ldnt(vp + j * kSliceRows);
				if (j0 + k >= w) c[k] = (I)-1;
			}
#pragma unroll
			for (int k = 0; k < kSliceUnroll; ++k) xv[k] = x(c[k] >= 0 ? c[k] : (I)0);
#pragma unroll
			for (int k = 0; k < kSliceUnroll; ++k) {
				const ACC t = sum + (ACC)v[k] * (ACC)xv[k];
				const bool isd = (int64_t)c[k] == r;
				if constexpr (is_offdiag<Epi>::value) {
					sum = c[k] >= 0 && !isd ? t : sum;
					dgv = isd ? (double)v[k] : dgv;
				}
				else
					sum = c[k] >= 0 ? t : sum;
				if (isd) {
					xself = xv[k];
					has_self = true;
				}
			}
		}
		if constexpr (is_offdiag<Epi>::value)
			epi.row_offdiag(r, sum, dgv, [&] { return has_self ? xself : x(r); });
		else
			epi.row(r, sum, d, [&] { return has_self ? xself : x(r); });
	}
	epi.block_end(d, blockIdx.x, gridDim.x);
}

// Slice-offset layout: wave = slice of 64 rows, lane = row.  The slice's
// offsets and lane masks are wave-uniform (scalar loads); values stream
// entry-major; the gathers x[row + off] of a wave are contiguous.  A lane
// skips the offsets its row does not have, so each row's sum runs over its
// entries in CSR (ascending column) order: bit-identical to the serial loop.
template<class Epi, class ACC, class V, class G, int BT>
__global__ __launch_bounds__(BT) void spmv_sdia_kernel(
	int64_t nrows, const int64_t * __restrict__ dg_ptr,
	const int32_t * __restrict__ dg_off, const uint64_t * __restrict__ dg_mask,
	const V * __restrict__ dg_val, G x, Epi epi) {
	if (epi.skip()) return;
	x.init();
	const int lane = threadIdx.x & (kSliceRows - 1);
	double d[Epi::NV];
#pragma unroll
	for (int i = 0; i < Epi::NV; ++i) d[i] = 0.0;
	constexpr int kWaves = BT / kSliceRows;
	const int64_t nslices = (nrows + kSliceRows - 1) / kSliceRows;
	const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kSliceRows));
	const int64_t sstride = (int64_t)gridDim.x * kWaves;
	for (int64_t s = (int64_t)blockIdx.x * kWaves + wave; s < nslices; s += sstride) {
		const int64_t r = s * kSliceRows + lane;
		const int64_t base = dg_ptr[s];
		const int w = (int)((dg_ptr[s + 1] - base) / kSliceRows);
		const int64_t q0 = base / kSliceRows;
		const V * vp = dg_val + base + lane;
		ACC sum = ACC(0);
		typename G::value_type xself{};
		bool has_self = false;
		double dgv = 0.; // diagonal value (offdiag epilogues)
		// Branch-free body: every load of a tile is issued before the first
		// use (slots past the width re-read the last column, absent lanes
		// gather x[0]); a select keeps `sum` unchanged for absent entries,
		// bitwise the same as skipping them.
		for (int j0 = 0; j0 < w; j0 += kSliceUnroll) {
			int32_t off[kSliceUnroll];
			bool pr[kSliceUnroll];
			V v[kSliceUnroll];
			typename G::value_type xv[kSliceUnroll];
#pragma unroll
			for (int k = 0; k < kSliceUnroll; ++k) {
				const int j = min(j0 + k, w - 1);
				// Offsets and masks: unconditional scalar loads of 8 consecutive
				// entries (the arrays carry kSliceUnroll entries of padding),
				// which merge into wide loads behind one wait.  A short-circuit
				// && or a clamped index here gives one load-and-wait per entry.
				off[k] = dg_off[q0 + j0 + k];
				const uint64_t m = dg_mask[q0 + j0 + k];
				pr[k] = (j0 + k < w) & (((m >> lane) & 1) != 0);
				v[k] = ldnt(vp + (int64_t)j * kSliceRows);
			}
#pragma unroll
			for (int k = 0; k < kSliceUnroll; ++k) xv[k] = x(pr[k] ? r + off[k] : 0);
#pragma unroll
			for (int k = 0; k < kSliceUnroll; ++k) {
				const ACC t = sum + (ACC)v[k] * (ACC)xv[k];
				const bool isd = pr[k] && off[k] == 0;
				if constexpr (is_offdiag<Epi>::value) {
					sum = pr[k] && !isd ? t : sum;
					dgv = isd ? (double)v[k] : dgv;
				}
				else
					sum = pr[k] ? t : sum;
				if (isd) {
					xself = xv[k];
					has_self = true;
				}
			}
		}
		if (r < nrows) {
			if constexpr (is_offdiag<Epi>::value)
				epi.row_offdiag(r, sum, dgv, [&] { return has_self ? xself : x(r); });
			else
				epi.row(r, sum, d, [&] { return has_self ? xself : x(r); });
		}
	}
	epi.block_end(d, blockIdx.x, gridDim.x);
}

// Symmetric stencil layout: one ascending offset list for the whole matrix,
// passed by value, so a wave loads no per-slice pointer or offsets -- only
// the slice's lane masks (w consecutive words, one scalar load).  Values are
// stored for the offsets >= 0 only; the entry of row r at offset o < 0 is
// the stored entry of row r + o at offset -o, read from that row's slot (a
// recently streamed slice: an L2 hit, not HBM traffic).  The layout is used
// only when the CSR values are bitwise symmetric, so each row is still
// summed over its own entries' values in CSR order: bit-identical to the
// serial loop.
struct stencil_desc {
	int w;  // offsets in the list (<= kStencilMax)
	int wu; // offsets >= 0: stored values per row
	int32_t off[kStencilMax];
	int32_t ju[kStencilMax]; // value column of entry j (of its mirror for off < 0)
};

inline stencil_desc make_stencil_desc(const csr_index & X) {
	stencil_desc sd{};
	sd.w = X.st_w;
	sd.wu = X.st_wu;
	for (int k = 0; k < kStencilMax; ++k) {
		sd.off[k] = X.st_off[k];
		sd.ju[k] = X.st_ju[k];
	}
	return sd;
}

// The slices are dealt to waves exactly as in the slice-offset kernel (same
// grid, same grid stride), so each block's partials cover the same rows and
// a whole solve runs bit for bit as on that layout.  Per slice, every value
// load is issued before the gathers and every gather before the first sum
// (scheduling barriers keep that order).  O32: value and gather addresses
// are the array base (uniform) plus a 32-bit byte offset per lane.
// MW: the minimum waves per SIMD the register allocation must allow (1: the
// compiler's own choice).  With 32-bit offsets every instance fits 8
// without spills.
template<class Epi, class ACC, class V, class G, int BT, bool O32, int MW>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(MW))) void
spmv_stencil_kernel(
	int64_t nrows, stencil_desc sd, const uint64_t * __restrict__ st_mask,
	const V * __restrict__ st_val, G x, Epi epi) {
	if (epi.skip()) return;
	x.init();
	using XV = typename G::value_type;
	const int lane = threadIdx.x & (kSliceRows - 1);
	double d[Epi::NV];
#pragma unroll
	for (int i = 0; i < Epi::NV; ++i) d[i] = 0.0;
	constexpr int kWaves = BT / kSliceRows;
	const int64_t nslices = (nrows + kSliceRows - 1) / kSliceRows;
	const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kSliceRows));
	const int64_t sstride = (int64_t)gridDim.x * kWaves;
	const int sw = sd.wu * kSliceRows; // values per slice
	for (int64_t s = (int64_t)blockIdx.x * kWaves + wave; s < nslices; s += sstride) {
		const int64_t r = s * kSliceRows + lane;
		const uint64_t * mp = st_mask + s * sd.w;
		V v[kStencilMax];
		XV xv[kStencilMax];
		// No load waits for the lane masks: a slot or vector element whose
		// row lies in [0, nrows) is read whether or not the entry exists
		// (the sum below skips absent entries), so every load of the slice
		// issues at once.  Entry k of row r sits in the slot of row q at
		// column ju[k]: q = r for offsets >= 0, the mirror row r + off[k]
		// for offsets < 0 (rows before the first read their own slot).
		// Row indices are 32-bit under O32 (nrows < 2^29 there).
		using RI = std::conditional_t<O32, int32_t, int64_t>;
		using RU = std::conditional_t<O32, uint32_t, uint64_t>;
		const RI ri = (RI)r;
#pragma unroll
		for (int k = 0; k < kStencilMax; ++k) {
			const RI rr = ri + (RI)sd.off[k];
			const bool inb = (RU)rr < (RU)nrows;
			const RI q = (sd.off[k] < 0 && inb) ? rr : ri;
			const RU e = (RU)(q >> 6) * (RU)sw + (RU)(sd.ju[k] * kSliceRows) + (RU)(q & 63);
			if constexpr (O32)
				v[k] = ld32(st_val, e);
			else
				v[k] = st_val[e];
		}
		__builtin_amdgcn_sched_barrier(0);
#pragma unroll
		for (int k = 0; k < kStencilMax; ++k) {
			const RI rr = ri + (RI)sd.off[k];
			const bool inb = (RU)rr < (RU)nrows;
			if constexpr (O32)
				xv[k] = x.at32(inb ? (uint32_t)rr : 0u);
			else
				xv[k] = x(inb ? (int64_t)rr : 0);
		}
		__builtin_amdgcn_sched_barrier(0);
		bool pr[kStencilMax];
#pragma unroll
		for (int k = 0; k < kStencilMax; ++k) {
			// the mask array carries kStencilMax words of padding; entries
			// past w have offset 0 and column 0 and are never present
			const uint64_t m = mp[k];
			pr[k] = (k < sd.w) & (((m >> lane) & 1) != 0);
		}
		ACC sum = ACC(0);
		XV xself{};
		bool has_self = false;
		double dgv = 0.; // diagonal value (offdiag epilogues)
#pragma unroll
		for (int k = 0; k < kStencilMax; ++k) {
			const ACC t = sum + (ACC)v[k] * (ACC)xv[k];
			const bool isd = pr[k] && sd.off[k] == 0;
			if constexpr (is_offdiag<Epi>::value) {
				sum = pr[k] && !isd ? t : sum;
				dgv = isd ? (double)v[k] : dgv;
			}
			else
				sum = pr[k] ? t : sum;
			if (isd) {
				xself = xv[k];
				has_self = true;
			}
		}
		if (r < nrows) {
			if constexpr (is_offdiag<Epi>::value)
				epi.row_offdiag(r, sum, dgv, [&] { return has_self ? xself : x(r); });
			else
				epi.row(r, sum, d, [&] { return has_self ? xself : x(r); });
		}
	}
	epi.block_end(d, blockIdx.x, gridDim.x);
}

// Epilogues that leave per-block partials cap the grid (fewer partials for
// the finalize kernel); plain products use one row per thread.
inline unsigned sliced_blocks(int64_t nrows, bool reduces = true) {
	const int64_t g = (nrows + kSpmvThreads - 1) / kSpmvThreads;
	return (unsigned)(reduces ? std::min<int64_t>(g, kSlicedMaxBlocks) : g);
}

// Slice-offset launches.  Plain products: 256-thread blocks, one slice per
// wave.  Reducing epilogues leave one partial per block.  Measured on CG
// iterations (256^3 and BiCGSTAB on 128^3):
//  * FP64 values: one slice per wave; 512-thread blocks on the 16.7M-row
//    matrix (0.689 s CG vs 0.705 s with 256 threads), 256-thread blocks on
//    the 2.1M-row one (28.3 ms BiCGSTAB vs 30.5 ms with 512).  The switch is
//    at kSdiaWideRows.
//  * FP32 values, whose slices are half the bytes: 256-thread blocks capped
//    at 8192 and grid-striding (ahead of one slice per wave; caps of 4096
//    and 512-thread blocks were slower, 16384 equal).
// Measured alternative, not used: one partial per wave (no end-of-block
// barrier, no cap) reduced in two fixed-order stages -- the extra stage
// cost more than the barrier saved (FP64 within 1%, FP32 slower).
constexpr int64_t kSdiaWideRows = (int64_t)1 << 23;

struct sdia_reduce_shape {
	int threads;
	unsigned max_blocks;
};

inline sdia_reduce_shape sdia_reduce(fsx_dtype vt, int64_t nrows) {
	if (vt == FSX_F64) return {nrows >= kSdiaWideRows ? 512 : 256, 1u << 17};
	return {256, 8192u};
}

inline unsigned sdia_blocks(int64_t nrows, bool reduces, int reduce_threads,
                            unsigned max_blocks) {
	const int64_t ns = (nrows + kSliceRows - 1) / kSliceRows;
	if (!reduces) return (unsigned)((ns * kSliceRows + kSpmvThreads - 1) / kSpmvThreads);
	const int64_t per = reduce_threads / kSliceRows;
	return (unsigned)std::min<int64_t>((ns + per - 1) / per, max_blocks);
}

// Launch the selected mapping for matrix A with gather policy G.
template<class Epi, class ACC, class V, class I, class G>
void spmv_launch_gather(hipStream_t st, const fsx_csr * A, const G & x,
                        const Epi & epi) {
	const I * rp = (const I *)A->rowptr();
	const I * ci = (const I *)A->colind();
	const V * vv = (const V *)A->vals->p;
	if (A->nrows == 0) return;
	if (A->plan_kind == 4) {
		const csr_index & X = *A->idx;
		const stencil_desc sd = make_stencil_desc(X);
		// 32-bit offsets when the value array and the gathered vector (at
		// most 8-byte elements) span less than 4 GiB
		const bool o32 = A->st_vals->bytes < ((size_t)1 << 32) &&
		                 (uint64_t)A->ncols * 8 < ((uint64_t)1 << 32) &&
		                 (uint64_t)A->nrows * 8 < ((uint64_t)1 << 32);
		// Reducing products at 8 waves per SIMD (FP64 CG K1 at 256^3: 0.198
		// against 0.215 ms), plain products at the compiler's choice (FP64
		// y = A x: 0.196 against 0.229 ms)
		constexpr int MW = Epi::reduces ? 8 : 1;
		auto go = [&](auto bt_c, unsigned g) {
			constexpr int BT = decltype(bt_c)::value;
			if (o32)
				hipLaunchKernelGGL((spmv_stencil_kernel<Epi, ACC, V, G, BT, true, MW>), dim3(g),
				                   dim3(BT), 0, st, A->nrows, sd, (const uint64_t *)X.st_mask->p,
				                   (const V *)A->st_vals->p, x, epi);
			else
				hipLaunchKernelGGL((spmv_stencil_kernel<Epi, ACC, V, G, BT, false, 1>), dim3(g),
				                   dim3(BT), 0, st, A->nrows, sd, (const uint64_t *)X.st_mask->p,
				                   (const V *)A->st_vals->p, x, epi);
		};
		if constexpr (!Epi::reduces)
			go(std::integral_constant<int, kSpmvThreads>{},
			   sdia_blocks(A->nrows, false, 0, 0));
		else {
			const sdia_reduce_shape sh = sdia_reduce(A->vt, A->nrows);
			const unsigned g = sdia_blocks(A->nrows, true, sh.threads, sh.max_blocks);
			if (sh.threads == 512) go(std::integral_constant<int, 512>{}, g);
			else go(std::integral_constant<int, 256>{}, g);
		}
	}
	else if (A->plan_kind == 3) {
		FSX_REQUIRE(!A->dg_stale, "slice-offset values lag the CSR values");
		auto go = [&](auto bt_c, unsigned g) {
			constexpr int BT = decltype(bt_c)::value;
			hipLaunchKernelGGL((spmv_sdia_kernel<Epi, ACC, V, G, BT>), dim3(g), dim3(BT), 0,
			                   st, A->nrows, (const int64_t *)A->idx->dg_ptr->p,
			                   (const int32_t *)A->idx->dg_off->p,
			                   (const uint64_t *)A->idx->dg_mask->p,
			                   (const V *)A->dg_vals->p, x, epi);
		};
		if constexpr (!Epi::reduces)
			go(std::integral_constant<int, kSpmvThreads>{},
			   sdia_blocks(A->nrows, false, 0, 0));
		else {
			const sdia_reduce_shape sh = sdia_reduce(A->vt, A->nrows);
			const unsigned g = sdia_blocks(A->nrows, true, sh.threads, sh.max_blocks);
			if (sh.threads == 512) go(std::integral_constant<int, 512>{}, g);
			else go(std::integral_constant<int, 256>{}, g);
		}
	}
	else if (A->plan_kind == 2) {
		const unsigned g = sliced_blocks(A->nrows, Epi::reduces);
		hipLaunchKernelGGL((spmv_sliced_kernel<Epi, ACC, V, I, G>), dim3(g),
		                   dim3(kSpmvThreads), 0, st, A->nrows,
		                   (const I *)A->idx->sl_ptr->p, (const I *)A->idx->sl_col->p,
		                   (const V *)A->sl_vals->p, x, epi);
	}
	else if (A->plan_kind == 0) {
		unsigned g = (unsigned)((A->nrows + kStreamRows - 1) / kStreamRows);
		hipLaunchKernelGGL((spmv_stream_kernel<Epi, ACC, V, I, G>), dim3(g),
		                   dim3(kSpmvThreads), 0, st, A->nrows, rp, ci, vv, x, epi);
	}
	else {
		auto go = [&](auto lanes_c) {
			constexpr int L = decltype(lanes_c)::value;
			constexpr int RPB = kSpmvThreads / L;
			unsigned g = (unsigned)((A->nrows + RPB - 1) / RPB);
			hipLaunchKernelGGL((spmv_vector_kernel<L, Epi, ACC, V, I, G>), dim3(g),
			                   dim3(kSpmvThreads), 0, st, A->nrows, rp, ci, vv, x,
			                   epi);
		};
		switch (A->plan_lanes) {
		case 2: go(std::integral_constant<int, 2>{}); break;
		case 4: go(std::integral_constant<int, 4>{}); break;
		case 8: go(std::integral_constant<int, 8>{}); break;
		case 16: go(std::integral_constant<int, 16>{}); break;
		case 32: go(std::integral_constant<int, 32>{}); break;
		default: go(std::integral_constant<int, 64>{}); break;
		}
	}
	FSX_HIP(hipGetLastError());
}

// y = A x with x a plain vector of element type X.
template<class Epi, class ACC, class V, class I, class X>
void spmv_launch_typed(hipStream_t st, const fsx_csr * A, const X * x,
                       const Epi & epi) {
	spmv_launch_gather<Epi, ACC, V, I>(st, A, gather_x<X>{x}, epi);
}

// Number of partials (each NV doubles) a reducing product on A leaves.
inline unsigned spmv_blocks(const fsx_csr * A) {
	if (A->plan_kind == 3 || A->plan_kind == 4) {
		const sdia_reduce_shape sh = sdia_reduce(A->vt, A->nrows);
		return sdia_blocks(A->nrows, true, sh.threads, sh.max_blocks);
	}
	if (A->plan_kind == 2) return sliced_blocks(A->nrows);
	if (A->plan_kind == 0)
		return (unsigned)((A->nrows + kStreamRows - 1) / kStreamRows);
	const int rpb = kSpmvThreads / A->plan_lanes;
	return (unsigned)((A->nrows + rpb - 1) / rpb);
}

// Doubles of partials space a reducing product on A needs.
inline size_t spmv_partials_size(const fsx_csr * A, int NV) {
	return (size_t)spmv_blocks(A) * NV;
}

// Sum the partials a reducing product on A left in `part` and run f on the
// totals (thread 0 of a one-block kernel).
template<int NV, class F>
void spmv_finalize(hipStream_t s, const fsx_csr * A, double * part, const F & f) {
	hipLaunchKernelGGL((finalize_kernel<NV, F>), dim3(1), dim3(kFinThreads), 0, s,
	                   (const double *)part, spmv_blocks(A), f);
}

// Dispatch on index and value type.  `body` is called with (V*, I*) tags.
template<class B>
void by_matrix_type(const fsx_csr * A, B && body) {
	if (A->vt == FSX_F64) {
		if (A->it == FSX_I32) body((double *)nullptr, (int32_t *)nullptr);
		else body((double *)nullptr, (int64_t *)nullptr);
	}
	else {
		if (A->it == FSX_I32) body((float *)nullptr, (int32_t *)nullptr);
		else body((float *)nullptr, (int64_t *)nullptr);
	}
}

} // namespace fsx
