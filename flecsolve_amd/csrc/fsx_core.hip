// fsx_core.hip -- context, device vectors, BLAS-1 kernels and deterministic
// reductions of libfsx.
//
// The reference's vector operations are FleCSI tasks, one per operation
// (flecsolve/vectors/operations/topo_tasks.hh:52-323), dispatched on field
// aliasing (flecsolve/vectors/operations/topo_view.hh:42-285); their
// serial semantics are vectors/seq.hh:93-304.  Here each operation is one
// HBM-streaming kernel.  Aliasing needs no dispatch: every kernel reads an
// element before writing it from the same lane, so any output may alias any
// input.  Arithmetic follows the reference's expression order and the
// library is compiled without FP contraction, so elementwise results are
// bit-identical to the reference's serial loops.
#include "fsx_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <random>

namespace fsx {

static thread_local std::string g_last_error;
void set_error(const std::string & msg) { g_last_error = msg; }
const char * get_error() { return g_last_error.c_str(); }

namespace ledger {
// base address -> sizes of the live buffers starting there (wrapped caller
// memory may be registered more than once)
static std::mutex g_mu;
static std::multimap<uintptr_t, size_t> g_live;

void add(const void * p, size_t bytes) {
	if (!p || !bytes) return;
	std::lock_guard<std::mutex> lk(g_mu);
	g_live.emplace((uintptr_t)p, bytes);
}
void remove(const void * p, size_t bytes) {
	if (!p || !bytes) return;
	std::lock_guard<std::mutex> lk(g_mu);
	auto rg = g_live.equal_range((uintptr_t)p);
	for (auto it = rg.first; it != rg.second; ++it)
		if (it->second == bytes) {
			g_live.erase(it);
			return;
		}
}
bool live(const void * p, size_t bytes) {
	const uintptr_t a = (uintptr_t)p;
	std::lock_guard<std::mutex> lk(g_mu);
	// buffers starting at or below a, nearest first
	for (auto it = g_live.upper_bound(a); it != g_live.begin();) {
		--it;
		if (a + bytes <= it->first + it->second) return true;
	}
	return false;
}
} // namespace ledger

// hipMalloc calls made and bytes held by live owned allocations
static std::atomic<uint64_t> g_alloc_count{0}, g_alloc_live{0};

dbuf::dbuf(size_t b) : bytes(b) {
	if (b == 0) return;
	hipError_t e = hipMalloc(&p, b);
	g_alloc_count.fetch_add(1);
	if (e != hipSuccess) {
		p = nullptr;
		throw error(FSX_ERR_ALLOC, std::string("hipMalloc(") +
		                               std::to_string(b) + " B): " +
		                               hipGetErrorString(e));
	}
	g_alloc_live.fetch_add(b);
	ledger::add(p, b);
}
void dbuf::adopt(void * q, size_t b) {
	ledger::remove(p, bytes);
	p = q;
	bytes = b;
	owns = false;
	ledger::add(p, bytes);
}
dbuf::~dbuf() {
	ledger::remove(p, bytes);
	if (p && owns) {
		(void)hipFree(p);
		g_alloc_live.fetch_sub(bytes);
	}
}

void check_vec(const fsx_vec * v, const char * name) {
	if (!v) throw error(FSX_ERR_ARG, std::string("null vector: ") + name);
}
void check_same(const fsx_vec * a, const fsx_vec * b) {
	if (a->n != b->n)
		throw error(FSX_ERR_ARG, "vector sizes differ (" + std::to_string(a->n) +
		                             " vs " + std::to_string(b->n) + ")");
	if (a->dt != b->dt) throw error(FSX_ERR_ARG, "vector dtypes differ");
}

static void check_ctx(const fsx_ctx * c) {
	if (!c) throw error(FSX_ERR_ARG, "null context");
}

// ---- elementwise kernels -------------------------------------------------

constexpr int kEwThreads = 256;

template<class T, class F>
__global__ __launch_bounds__(kEwThreads) void ew_kernel(size_t n, F f) {
	size_t i = (size_t)blockIdx.x * kEwThreads + threadIdx.x;
	size_t stride = (size_t)gridDim.x * kEwThreads;
	for (; i < n; i += stride) f(i);
}

static unsigned ew_grid(size_t n) {
	size_t g = (n + kEwThreads - 1) / kEwThreads;
	if (g > 65536) g = 65536;
	if (g == 0) g = 1;
	return (unsigned)g;
}

template<class T, class F>
static void ew_launch(fsx_ctx * ctx, size_t n, F f) {
	if (n == 0) return;
	hipLaunchKernelGGL((ew_kernel<T, F>), dim3(ew_grid(n)), dim3(kEwThreads), 0,
	                   ctx->stream, n, f);
	FSX_HIP(hipGetLastError());
	ctx->launches++;
}

// Dispatch on dtype: body(T*) is instantiated for double and float.
template<class B>
static void by_type(const fsx_vec * v, B && body) {
	if (v->dt == FSX_F64) body((double *)nullptr);
	else body((float *)nullptr);
}

void launch_copy(fsx_ctx * ctx, const fsx_vec * s, fsx_vec * d) {
	check_same(s, d);
	if (s->ptr() == d->ptr()) return;
	FSX_HIP(hipMemcpyAsync(d->ptr(), s->ptr(), s->n * s->elem(),
	                       hipMemcpyDeviceToDevice, ctx->stream));
}

void launch_set(fsx_ctx * ctx, double a, fsx_vec * x) {
	by_type(x, [&](auto * tag) {
		using T = std::remove_pointer_t<decltype(tag)>;
		T * p = (T *)x->ptr();
		T v = (T)a;
		ew_launch<T>(ctx, x->n, [=] __device__(size_t i) { p[i] = v; });
	});
}

// z = a * x + y  (seq.hh:171-174)
void launch_axpy(fsx_ctx * ctx, double a, const fsx_vec * x, const fsx_vec * y,
                 fsx_vec * z) {
	check_same(x, y);
	check_same(x, z);
	by_type(x, [&](auto * tag) {
		using T = std::remove_pointer_t<decltype(tag)>;
		const T * xp = (const T *)x->ptr();
		const T * yp = (const T *)y->ptr();
		T * zp = (T *)z->ptr();
		T av = (T)a;
		ew_launch<T>(ctx, x->n,
		             [=] __device__(size_t i) { zp[i] = av * xp[i] + yp[i]; });
	});
}

// dst = scale * src with a precision change (FP64 <-> FP32); same dtype is a
// scaled copy.
void launch_convert(fsx_ctx * ctx, const fsx_vec * s, fsx_vec * d,
                    double scale) {
	FSX_REQUIRE(s->n == d->n, "convert: sizes differ");
	size_t n = s->n;
	if (s->dt == FSX_F64 && d->dt == FSX_F32) {
		const double * sp = (const double *)s->ptr();
		float * dp = (float *)d->ptr();
		ew_launch<float>(ctx, n, [=] __device__(size_t i) {
			dp[i] = (float)(scale * sp[i]);
		});
	}
	else if (s->dt == FSX_F32 && d->dt == FSX_F64) {
		const float * sp = (const float *)s->ptr();
		double * dp = (double *)d->ptr();
		ew_launch<double>(ctx, n, [=] __device__(size_t i) {
			dp[i] = scale * (double)sp[i];
		});
	}
	else {
		by_type(s, [&](auto * tag) {
			using T = std::remove_pointer_t<decltype(tag)>;
			const T * sp = (const T *)s->ptr();
			T * dp = (T *)d->ptr();
			T sc = (T)scale;
			ew_launch<T>(ctx, n, [=] __device__(size_t i) { dp[i] = sc * sp[i]; });
		});
	}
}

// ---- reductions ----------------------------------------------------------
//
// Two stages: each block accumulates a grid-stride slice in FP64 registers,
// reduces it over the block and stores a partial; a one-block kernel then
// combines the partials in block order.  The grid size is a fixed function of
// n, so the summation order -- and the result -- is the same on every run.  Reference: topo_tasks.hh:52-66 (scalar_prod),
// 216-297 (norms, min, max); folds in operations/topo_view.hh:212-274.

constexpr int kRedThreads = 256;

int red_grid(size_t n) {
	size_t g = (n + (size_t)kRedThreads * 8 - 1) / ((size_t)kRedThreads * 8);
	if (g < 1) g = 1;
	if (g > (size_t)kRedBlocks) g = kRedBlocks;
	return (int)g;
}

template<class T>
struct red_op;

template<red_kind K, class T>
__device__ __forceinline__ double red_elem(const T * x, const T * y, size_t i) {
	if constexpr (K == RED_DOT) return (double)x[i] * (double)y[i];
	else if constexpr (K == RED_L1) return fabs((double)x[i]);
	else if constexpr (K == RED_L2SQ) return (double)x[i] * (double)x[i];
	else if constexpr (K == RED_INF) return fabs((double)x[i]);
	else return (double)x[i];
}

template<red_kind K>
__device__ __forceinline__ double red_init() {
	if constexpr (K == RED_MIN) return __DBL_MAX__;
	else if constexpr (K == RED_MAX) return -__DBL_MAX__;
	// inf_norm starts from numeric_limits<double>::min() (vectors/seq.hh:251)
	else if constexpr (K == RED_INF) return __DBL_MIN__;
	else return 0.0;
}

template<red_kind K>
__device__ __forceinline__ double red_comb(double a, double b) {
	if constexpr (K == RED_MIN) return fmin(a, b);
	else if constexpr (K == RED_MAX || K == RED_INF) return fmax(a, b);
	else return a + b;
}

template<red_kind K>
__device__ __forceinline__ double wave_red(double v) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v = red_comb<K>(v, __shfl_xor(v, o, 64));
	return v;
}

template<red_kind K, class T>
__global__ __launch_bounds__(kRedThreads) void reduce_kernel(
	size_t n, const T * __restrict__ x, const T * __restrict__ y,
	double * __restrict__ partials) {
	double acc = red_init<K>();
	const size_t stride = (size_t)gridDim.x * kRedThreads;
	for (size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x; i < n;
	     i += stride)
		acc = red_comb<K>(acc, red_elem<K>(x, y, i));
	__shared__ double sh[4];
	acc = wave_red<K>(acc);
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
	__syncthreads();
	if (threadIdx.x == 0)
		partials[blockIdx.x] =
			red_comb<K>(red_comb<K>(red_comb<K>(sh[0], sh[1]), sh[2]), sh[3]);
}

// stage 2: one block combines the partials in a fixed order
template<red_kind K>
__global__ __launch_bounds__(kRedThreads) void reduce_finalize_kernel(
	const double * __restrict__ partials, unsigned nblk, double * out) {
	double v = red_init<K>();
	for (unsigned b = threadIdx.x; b < nblk; b += kRedThreads)
		v = red_comb<K>(v, partials[b]);
	__shared__ double sh[4];
	v = wave_red<K>(v);
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
	__syncthreads();
	if (threadIdx.x == 0)
		*out = red_comb<K>(red_comb<K>(red_comb<K>(sh[0], sh[1]), sh[2]), sh[3]);
}

// sum of pow(x_i, p): the general-p lp_norm of the FleCSI vector path
// (vectors/operations/topo_tasks.hh:216-226 takes std::pow(u, p) with no
// absolute value), folded in `real` = double for FP64 vectors
template<class T>
__global__ __launch_bounds__(kRedThreads) void powsum_kernel(size_t n, const T * __restrict__ x,
                                                             double p,
                                                             double * __restrict__ partials) {
	double acc = 0.0;
	const size_t stride = (size_t)gridDim.x * kRedThreads;
	for (size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += stride)
		acc += pow((double)x[i], p);
	__shared__ double sh[4];
	acc = wave_red<RED_DOT>(acc);
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
	__syncthreads();
	if (threadIdx.x == 0) partials[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// FP32 vectors: lp_norm_local folds into `real`, float here -- each term
// std::pow(float, int) (evaluated in double) is rounded to float and summed
// in float, within a block and then over the blocks in block order
__global__ __launch_bounds__(kRedThreads) void powsum_f32_kernel(size_t n,
                                                                 const float * __restrict__ x,
                                                                 double p,
                                                                 double * __restrict__ partials) {
	float acc = 0.f;
	const size_t stride = (size_t)gridDim.x * kRedThreads;
	for (size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += stride)
		acc += (float)pow((double)x[i], p);
	__shared__ float sh[4];
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
	__syncthreads();
	if (threadIdx.x == 0) partials[blockIdx.x] = (double)(((sh[0] + sh[1]) + sh[2]) + sh[3]);
}

__global__ __launch_bounds__(kRedThreads) void powsum_f32_finalize_kernel(
	const double * __restrict__ partials, unsigned nblk, double * out) {
	float v = 0.f;
	for (unsigned b = threadIdx.x; b < nblk; b += kRedThreads) v += (float)partials[b];
	__shared__ float sh[4];
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
	__syncthreads();
	if (threadIdx.x == 0) *out = (double)(((sh[0] + sh[1]) + sh[2]) + sh[3]);
}

template<red_kind K>
static void reduce_launch_k(fsx_ctx * ctx, fsx_dtype dt, size_t n, const void * x,
                            const void * y, double * dslot, double * part) {
	int g = red_grid(n);
	if (dt == FSX_F64)
		hipLaunchKernelGGL((reduce_kernel<K, double>), dim3(g), dim3(kRedThreads),
		                   0, ctx->stream, n, (const double *)x, (const double *)y, part);
	else
		hipLaunchKernelGGL((reduce_kernel<K, float>), dim3(g), dim3(kRedThreads),
		                   0, ctx->stream, n, (const float *)x, (const float *)y, part);
	hipLaunchKernelGGL((reduce_finalize_kernel<K>), dim3(1), dim3(kRedThreads), 0,
	                   ctx->stream, (const double *)part, (unsigned)g, dslot);
	FSX_HIP(hipGetLastError());
	ctx->launches += 2;
}

void reduce_raw(fsx_ctx * ctx, red_kind k, fsx_dtype dt, size_t n, const void * x,
                const void * y, double * dslot, double * part) {
	switch (k) {
	case RED_DOT: reduce_launch_k<RED_DOT>(ctx, dt, n, x, y, dslot, part); break;
	case RED_L1: reduce_launch_k<RED_L1>(ctx, dt, n, x, y, dslot, part); break;
	case RED_L2SQ: reduce_launch_k<RED_L2SQ>(ctx, dt, n, x, y, dslot, part); break;
	case RED_INF: reduce_launch_k<RED_INF>(ctx, dt, n, x, y, dslot, part); break;
	case RED_MIN: reduce_launch_k<RED_MIN>(ctx, dt, n, x, y, dslot, part); break;
	case RED_MAX: reduce_launch_k<RED_MAX>(ctx, dt, n, x, y, dslot, part); break;
	}
}

void reduce_async(fsx_ctx * ctx, red_kind k, const fsx_vec * x,
                  const fsx_vec * y, double * dslot) {
	check_vec(x, "x");
	if (y) check_same(x, y);
	double * part = ensure_partials(ctx, (size_t)red_grid(x->n));
	reduce_raw(ctx, k, x->dt, x->n, x->ptr(), y ? y->ptr() : nullptr, dslot, part);
}

void ctx_retain(fsx_ctx * c) { c->refs.fetch_add(1); }
void ctx_release(fsx_ctx * c) {
	if (c->refs.fetch_sub(1) == 1) {
		(void)hipSetDevice(c->device);
		(void)hipStreamSynchronize(c->stream);
		delete c;
	}
}

double * ensure_relax_buffer(fsx_ctx * ctx, size_t n) {
	if (n <= ctx->relax_cap) return (double *)ctx->relax_buf->p;
	FSX_HIP(hipStreamSynchronize(ctx->stream));
	ctx->relax_buf = std::make_unique<dbuf>(std::max<size_t>(n, 1) * sizeof(double));
	ctx->relax_cap = std::max<size_t>(n, 1);
	return (double *)ctx->relax_buf->p;
}

double * ensure_partials(fsx_ctx * ctx, size_t count) {
	if (count <= ctx->part_cap) return ctx->ws.partials;
	size_t cap = std::max<size_t>(count, 2 * ctx->part_cap);
	FSX_HIP(hipStreamSynchronize(ctx->stream));
	invalidate_graphs(ctx);
	ctx->part_buf = std::make_unique<dbuf>(cap * sizeof(double));
	ctx->part_cap = cap;
	ctx->ws.partials = (double *)ctx->part_buf->p;
	return ctx->ws.partials;
}

double fetch_scalar(fsx_ctx * ctx, const double * dslot) {
	FSX_HIP(hipMemcpyAsync(ctx->host_scalars, dslot, sizeof(double),
	                       hipMemcpyDeviceToHost, ctx->stream));
	FSX_HIP(hipStreamSynchronize(ctx->stream));
	ctx->host_syncs++;
	return ctx->host_scalars[0];
}

double reduce_host(fsx_ctx * ctx, red_kind k, const fsx_vec * x,
                   const fsx_vec * y) {
	if (x->n == 0) {
		switch (k) {
		case RED_MIN: return std::numeric_limits<double>::max();
		case RED_MAX: return std::numeric_limits<double>::lowest();
		case RED_INF: return std::numeric_limits<double>::min();
		default: return 0.0;
		}
	}
	double * slot = ctx->ws.scalars;
	reduce_async(ctx, k, x, y, slot);
	return fetch_scalar(ctx, slot);
}

} // namespace fsx

fsx_ctx::~fsx_ctx() {
	if (host_scalars) (void)hipHostFree(host_scalars);
	if (stream) (void)hipStreamDestroy(stream);
}

using namespace fsx;

// ---- C ABI ---------------------------------------------------------------

extern "C" {

const char * fsx_last_error(void) { return get_error(); }

int fsx_reduce_async(fsx_ctx * ctx, int kind, const fsx_vec * x, const fsx_vec * y,
                     double * d_out) {
	return guard([&] {
		FSX_REQUIRE(ctx && d_out, "null argument");
		FSX_REQUIRE(kind >= FSX_RED_DOT && kind <= FSX_RED_MAX, "unknown reduction");
		FSX_REQUIRE(kind != FSX_RED_DOT || y, "dot needs y");
		reduce_async(ctx, (red_kind)kind, x, kind == FSX_RED_DOT ? y : nullptr, d_out);
	});
}

int fsx_ctx_create(int device, fsx_ctx ** out) {
	return guard([&] {
		FSX_REQUIRE(out, "null output");
		int ndev = 0;
		FSX_HIP(hipGetDeviceCount(&ndev));
		FSX_REQUIRE(device >= 0 && device < ndev,
		            "device " + std::to_string(device) + " not present (" +
		                std::to_string(ndev) + " devices)");
		FSX_HIP(hipSetDevice(device));
		auto c = std::make_unique<fsx_ctx>();
		c->device = device;
		hipDeviceProp_t prop;
		FSX_HIP(hipGetDeviceProperties(&prop, device));
		c->num_cus = prop.multiProcessorCount;
		// default-priority, non-blocking stream
		FSX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
		// the no-allocation CG's scratch address on this device
		noalloc_init(device);
		size_t ws_bytes = sizeof(double) * kScalarSlots;
		c->ws_buf = std::make_unique<dbuf>(ws_bytes);
		char * base = (char *)c->ws_buf->p;
		c->ws.scalars = (double *)base;
		ensure_partials(c.get(), (size_t)kRedBlocks * 4);
		FSX_HIP(hipMemsetAsync(base, 0, ws_bytes, c->stream));
		FSX_HIP(hipHostMalloc((void **)&c->host_scalars,
		                      sizeof(double) * kScalarSlots));
		FSX_HIP(hipStreamSynchronize(c->stream));
		*out = c.release();
	});
}

int fsx_ctx_destroy(fsx_ctx * ctx) {
	return guard([&] {
		if (!ctx) return;
		(void)hipStreamSynchronize(ctx->stream);
		// pooled solver work vectors reference the context: drop them first
		ctx->cg_cache.reset();
		ctx_release(ctx);
	});
}

int fsx_sync(fsx_ctx * ctx) {
	return guard([&] {
		check_ctx(ctx);
		FSX_HIP(hipStreamSynchronize(ctx->stream));
	});
}

int fsx_alloc_stats(uint64_t * allocations, uint64_t * live_bytes) {
	return guard([&] {
		if (allocations) *allocations = g_alloc_count.load();
		if (live_bytes) *live_bytes = g_alloc_live.load();
	});
}

int fsx_ctx_stream(fsx_ctx * ctx, void ** stream) {
	return guard([&] {
		check_ctx(ctx);
		*stream = (void *)ctx->stream;
	});
}

int fsx_ctx_counters(fsx_ctx * ctx, uint64_t * host_syncs, uint64_t * launches) {
	return guard([&] {
		check_ctx(ctx);
		if (host_syncs) *host_syncs = ctx->host_syncs;
		if (launches) *launches = ctx->launches;
	});
}

int fsx_vec_create(fsx_ctx * ctx, size_t n, fsx_dtype dt, fsx_vec ** out) {
	return guard([&] {
		check_ctx(ctx);
		FSX_REQUIRE(out, "null output");
		FSX_REQUIRE(dt == FSX_F64 || dt == FSX_F32, "bad dtype");
		auto v = std::make_unique<fsx_vec>();
		v->ctx = ctx;
		v->n = n;
		v->dt = dt;
		v->buf = std::make_shared<dbuf>(n * v->elem());
		v->id = ctx->next_id++;
		ctx_retain(ctx);
		if (n) {
			// value-initialised like std::vector<double>(n) in seq_work
			FSX_HIP(hipMemsetAsync(v->buf->p, 0, n * v->elem(), ctx->stream));
		}
		*out = v.release();
	});
}

int fsx_vec_wrap(fsx_ctx * ctx, void * dptr, size_t n, fsx_dtype dt,
                 fsx_vec ** out) {
	return guard([&] {
		check_ctx(ctx);
		FSX_REQUIRE(out, "null output");
		FSX_REQUIRE(dptr || n == 0, "null device pointer");
		auto v = std::make_unique<fsx_vec>();
		v->ctx = ctx;
		v->n = n;
		v->dt = dt;
		v->buf = std::make_shared<dbuf>();
		v->buf->adopt(dptr, n * v->elem());
		v->id = ctx->next_id++;
		ctx_retain(ctx);
		*out = v.release();
	});
}

int fsx_vec_retain(fsx_vec * v) {
	return guard([&] {
		check_vec(v, "v");
		v->refcnt.fetch_add(1);
	});
}

int fsx_vec_release(fsx_vec * v) {
	return guard([&] {
		if (!v) return;
		if (v->refcnt.fetch_sub(1) == 1) {
			// pending kernels may still read the buffer
			fsx_ctx * c = v->ctx;
			(void)hipStreamSynchronize(c->stream);
			delete v;
			ctx_release(c);
		}
	});
}

int fsx_vec_size(const fsx_vec * v, size_t * n) {
	return guard([&] { check_vec(v, "v"); *n = v->n; });
}
int fsx_vec_dtype(const fsx_vec * v, int * dt) {
	return guard([&] { check_vec(v, "v"); *dt = (int)v->dt; });
}
int fsx_vec_id(const fsx_vec * v, uint64_t * id) {
	return guard([&] { check_vec(v, "v"); *id = v->id; });
}
int fsx_vec_device_ptr(const fsx_vec * v, void ** dptr) {
	return guard([&] { check_vec(v, "v"); *dptr = v->ptr(); });
}

int fsx_vec_upload(fsx_ctx * ctx, fsx_vec * v, const void * host, size_t n) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(v, "v");
		FSX_REQUIRE(n == v->n, "upload: size mismatch");
		if (n == 0) return;
		FSX_HIP(hipMemcpyAsync(v->ptr(), host, n * v->elem(),
		                       hipMemcpyHostToDevice, ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
	});
}

int fsx_vec_download(fsx_ctx * ctx, const fsx_vec * v, void * host, size_t n) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(v, "v");
		FSX_REQUIRE(n == v->n, "download: size mismatch");
		if (n == 0) return;
		FSX_HIP(hipMemcpyAsync(host, v->ptr(), n * v->elem(),
		                       hipMemcpyDeviceToHost, ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
	});
}

int fsx_vec_get(fsx_ctx * ctx, const fsx_vec * v, size_t i, double * out) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(v, "v");
		FSX_REQUIRE(out, "null output");
		FSX_REQUIRE(i < v->n, "element " + std::to_string(i) + " of a vector of " +
		                          std::to_string(v->n));
		double * h = ctx->host_scalars;
		FSX_HIP(hipMemcpyAsync(h, (const char *)v->ptr() + i * v->elem(), v->elem(),
		                       hipMemcpyDeviceToHost, ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		ctx->host_syncs++;
		*out = v->dt == FSX_F64 ? h[0] : (double)*(const float *)h;
	});
}

int fsx_vec_set(fsx_ctx * ctx, fsx_vec * v, size_t i, double value) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(v, "v");
		FSX_REQUIRE(i < v->n, "element " + std::to_string(i) + " of a vector of " +
		                          std::to_string(v->n));
		// staged in the context's pinned scalars: the copy has completed
		// when the call returns
		double * h = ctx->host_scalars;
		if (v->dt == FSX_F64) h[0] = value;
		else *(float *)h = (float)value;
		FSX_HIP(hipMemcpyAsync((char *)v->ptr() + i * v->elem(), h, v->elem(),
		                       hipMemcpyHostToDevice, ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
	});
}

// ---- BLAS-1 --------------------------------------------------------------

#define FSX_EW3(NAME, EXPR)                                                 \
	int NAME(fsx_ctx * ctx, const fsx_vec * x, const fsx_vec * y,           \
	         fsx_vec * z) {                                                 \
		return guard([&] {                                                  \
			check_ctx(ctx);                                                 \
			check_vec(x, "x"); check_vec(y, "y"); check_vec(z, "z");        \
			check_same(x, y); check_same(x, z);                             \
			by_type(x, [&](auto * tag) {                                    \
				using T = std::remove_pointer_t<decltype(tag)>;             \
				const T * xp = (const T *)x->ptr();                         \
				const T * yp = (const T *)y->ptr();                         \
				T * zp = (T *)z->ptr();                                     \
				ew_launch<T>(ctx, x->n, [=] __device__(size_t i) {          \
					const T a = xp[i], b = yp[i];                           \
					zp[i] = EXPR;                                           \
				});                                                         \
			});                                                             \
		});                                                                 \
	}

FSX_EW3(fsx_add, a + b)       // seq.hh:132-135
FSX_EW3(fsx_subtract, a - b)  // seq.hh:137-140
FSX_EW3(fsx_multiply, a * b)  // seq.hh:142-145
FSX_EW3(fsx_divide, a / b)    // seq.hh:147-150

int fsx_copy(fsx_ctx * ctx, const fsx_vec * src, fsx_vec * dst) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(src, "src"); check_vec(dst, "dst");
		launch_copy(ctx, src, dst);
	});
}

int fsx_zero(fsx_ctx * ctx, fsx_vec * x) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(x, "x");
		launch_set(ctx, 0.0, x);
	});
}

int fsx_set_scalar(fsx_ctx * ctx, double a, fsx_vec * x) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(x, "x");
		launch_set(ctx, a, x);
	});
}

int fsx_scale(fsx_ctx * ctx, double a, fsx_vec * x) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(x, "x");
		by_type(x, [&](auto * tag) {
			using T = std::remove_pointer_t<decltype(tag)>;
			T * p = (T *)x->ptr();
			T av = (T)a;
			ew_launch<T>(ctx, x->n, [=] __device__(size_t i) { p[i] *= av; });
		});
	});
}

int fsx_scale_to(fsx_ctx * ctx, double a, const fsx_vec * x, fsx_vec * y) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(x, "x"); check_vec(y, "y");
		check_same(x, y);
		// topo_view.hh:72-73 asserts the two fields differ
		FSX_REQUIRE(x->ptr() != y->ptr() || x->n == 0,
		            "scale operation: vector data cannot be the same");
		by_type(x, [&](auto * tag) {
			using T = std::remove_pointer_t<decltype(tag)>;
			const T * xp = (const T *)x->ptr();
			T * yp = (T *)y->ptr();
			T av = (T)a;
			ew_launch<T>(ctx, x->n,
			             [=] __device__(size_t i) { yp[i] = av * xp[i]; });
		});
	});
}

int fsx_reciprocal(fsx_ctx * ctx, const fsx_vec * x, fsx_vec * y) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(x, "x"); check_vec(y, "y");
		check_same(x, y);
		by_type(x, [&](auto * tag) {
			using T = std::remove_pointer_t<decltype(tag)>;
			const T * xp = (const T *)x->ptr();
			T * yp = (T *)y->ptr();
			ew_launch<T>(ctx, x->n,
			             [=] __device__(size_t i) { yp[i] = T(1.0) / xp[i]; });
		});
	});
}

int fsx_linear_sum(fsx_ctx * ctx, double a, const fsx_vec * x, double b,
                   const fsx_vec * y, fsx_vec * z) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(x, "x"); check_vec(y, "y"); check_vec(z, "z");
		check_same(x, y); check_same(x, z);
		by_type(x, [&](auto * tag) {
			using T = std::remove_pointer_t<decltype(tag)>;
			const T * xp = (const T *)x->ptr();
			const T * yp = (const T *)y->ptr();
			T * zp = (T *)z->ptr();
			T av = (T)a, bv = (T)b;
			ew_launch<T>(ctx, x->n, [=] __device__(size_t i) {
				zp[i] = av * xp[i] + bv * yp[i];
			});
		});
	});
}

int fsx_axpy(fsx_ctx * ctx, double a, const fsx_vec * x, const fsx_vec * y,
             fsx_vec * z) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(x, "x"); check_vec(y, "y"); check_vec(z, "z");
		launch_axpy(ctx, a, x, y, z);
	});
}

int fsx_axpby(fsx_ctx * ctx, double a, double b, const fsx_vec * x,
              fsx_vec * z) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(x, "x"); check_vec(z, "z");
		check_same(x, z);
		by_type(x, [&](auto * tag) {
			using T = std::remove_pointer_t<decltype(tag)>;
			const T * xp = (const T *)x->ptr();
			T * zp = (T *)z->ptr();
			T av = (T)a, bv = (T)b;
			ew_launch<T>(ctx, x->n, [=] __device__(size_t i) {
				zp[i] = av * xp[i] + bv * zp[i];
			});
		});
	});
}

int fsx_abs(fsx_ctx * ctx, const fsx_vec * x, fsx_vec * y) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(x, "x"); check_vec(y, "y");
		check_same(x, y);
		by_type(x, [&](auto * tag) {
			using T = std::remove_pointer_t<decltype(tag)>;
			const T * xp = (const T *)x->ptr();
			T * yp = (T *)y->ptr();
			ew_launch<T>(ctx, x->n,
			             [=] __device__(size_t i) { yp[i] = fabs(xp[i]); });
		});
	});
}

int fsx_add_scalar(fsx_ctx * ctx, const fsx_vec * x, double a, fsx_vec * y) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(x, "x"); check_vec(y, "y");
		check_same(x, y);
		by_type(x, [&](auto * tag) {
			using T = std::remove_pointer_t<decltype(tag)>;
			const T * xp = (const T *)x->ptr();
			T * yp = (T *)y->ptr();
			T av = (T)a;
			ew_launch<T>(ctx, x->n,
			             [=] __device__(size_t i) { yp[i] = xp[i] + av; });
		});
	});
}

// mt19937 is inherently sequential; values are drawn on the host exactly as
// seq_ops::set_random (vectors/seq.hh:103-115) and uploaded.
int fsx_set_random(fsx_ctx * ctx, fsx_vec * x, unsigned seed) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(x, "x");
		std::mt19937 gen(seed);
		std::uniform_real_distribution<double> dis(0., 1.);
		if (x->dt == FSX_F64) {
			std::vector<double> h(x->n);
			for (auto & v : h) v = dis(gen);
			FSX_HIP(hipMemcpyAsync(x->ptr(), h.data(), x->n * 8,
			                       hipMemcpyHostToDevice, ctx->stream));
			FSX_HIP(hipStreamSynchronize(ctx->stream));
		}
		else {
			std::vector<float> h(x->n);
			for (auto & v : h) v = (float)dis(gen);
			FSX_HIP(hipMemcpyAsync(x->ptr(), h.data(), x->n * 4,
			                       hipMemcpyHostToDevice, ctx->stream));
			FSX_HIP(hipStreamSynchronize(ctx->stream));
		}
	});
}

int fsx_convert(fsx_ctx * ctx, const fsx_vec * src, fsx_vec * dst,
                double scale) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(src, "src"); check_vec(dst, "dst");
		launch_convert(ctx, src, dst, scale);
	});
}

// ---- reductions ----------------------------------------------------------

int fsx_dot(fsx_ctx * ctx, const fsx_vec * x, const fsx_vec * y, double * out) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(x, "x"); check_vec(y, "y");
		*out = reduce_host(ctx, RED_DOT, x, y);
	});
}
int fsx_nrm1(fsx_ctx * ctx, const fsx_vec * x, double * out) {
	return guard([&] {
		check_ctx(ctx); check_vec(x, "x");
		*out = reduce_host(ctx, RED_L1, x, nullptr);
	});
}
int fsx_nrm2(fsx_ctx * ctx, const fsx_vec * x, double * out) {
	return guard([&] {
		check_ctx(ctx); check_vec(x, "x");
		*out = std::sqrt(reduce_host(ctx, RED_L2SQ, x, nullptr));
	});
}
int fsx_nrminf(fsx_ctx * ctx, const fsx_vec * x, double * out) {
	return guard([&] {
		check_ctx(ctx); check_vec(x, "x");
		*out = reduce_host(ctx, RED_INF, x, nullptr);
	});
}
int fsx_nrmp(fsx_ctx * ctx, const fsx_vec * x, int p, double * out) {
	return guard([&] {
		check_ctx(ctx); check_vec(x, "x");
		FSX_REQUIRE(out, "null output");
		FSX_REQUIRE(p >= 1, "lp norm: p must be >= 1");
		if (p == 1) { *out = reduce_host(ctx, RED_L1, x, nullptr); return; }
		if (p == 2) { *out = std::sqrt(reduce_host(ctx, RED_L2SQ, x, nullptr)); return; }
		double sum = 0.0;
		if (x->n > 0) {
			const int g = red_grid(x->n);
			double * part = ensure_partials(ctx, (size_t)g);
			double * slot = ctx->ws.scalars;
			if (x->dt == FSX_F64) {
				hipLaunchKernelGGL((powsum_kernel<double>), dim3(g), dim3(kRedThreads), 0,
				                   ctx->stream, x->n, (const double *)x->ptr(), (double)p, part);
				hipLaunchKernelGGL((reduce_finalize_kernel<RED_DOT>), dim3(1), dim3(kRedThreads),
				                   0, ctx->stream, (const double *)part, (unsigned)g, slot);
			}
			else {
				hipLaunchKernelGGL(powsum_f32_kernel, dim3(g), dim3(kRedThreads), 0, ctx->stream,
				                   x->n, (const float *)x->ptr(), (double)p, part);
				hipLaunchKernelGGL(powsum_f32_finalize_kernel, dim3(1), dim3(kRedThreads), 0,
				                   ctx->stream, (const double *)part, (unsigned)g, slot);
			}
			FSX_HIP(hipGetLastError());
			ctx->launches += 2;
			sum = fetch_scalar(ctx, slot);
		}
		*out = std::pow(sum, 1. / p);
	});
}
int fsx_min(fsx_ctx * ctx, const fsx_vec * x, double * out) {
	return guard([&] {
		check_ctx(ctx); check_vec(x, "x");
		*out = reduce_host(ctx, RED_MIN, x, nullptr);
	});
}
int fsx_max(fsx_ctx * ctx, const fsx_vec * x, double * out) {
	return guard([&] {
		check_ctx(ctx); check_vec(x, "x");
		*out = reduce_host(ctx, RED_MAX, x, nullptr);
	});
}

int fsx_jacobi_apply(fsx_ctx * ctx, const fsx_vec * dinv, const fsx_vec * r,
                     fsx_vec * z) {
	return guard([&] {
		check_ctx(ctx);
		check_vec(dinv, "dinv"); check_vec(r, "r"); check_vec(z, "z");
		check_same(dinv, r); check_same(r, z);
		by_type(r, [&](auto * tag) {
			using T = std::remove_pointer_t<decltype(tag)>;
			const T * dp = (const T *)dinv->ptr();
			const T * rp = (const T *)r->ptr();
			T * zp = (T *)z->ptr();
			// Dinv CSR SpMV row: y_i = 0; y_i += dinv_i * r_i
			ew_launch<T>(ctx, r->n,
			             [=] __device__(size_t i) { zp[i] = T(0) + dp[i] * rp[i]; });
		});
	});
}

} // extern "C"
