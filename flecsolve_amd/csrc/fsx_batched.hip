// fsx_batched.hip -- many independent CG solves in one persistent launch.
//
// Input: the block-diagonal concatenation of the systems (one fsx_csr, global
// column ids) and the row offsets of each system.  One launch, sized to the
// number of workgroups the GPU holds at once; each workgroup repeatedly takes
// the next system from a queue ordered largest first (longest-processing-time
// order) and runs the reference CG iteration (op::cg::apply,
// flecsolve/solvers/cg.hh:43-139, with z = D^{-1} r Jacobi or op::I) on that
// system alone.  Dot products are workgroup reductions in a fixed order (no
// grid-wide synchronisation), so each system's arithmetic -- and its
// iteration count -- does not depend on which workgroup solves it or when.
// Per-system iteration counts, stop reasons and final residuals are the
// convergence masks: a workgroup leaves a system the iteration it converges
// and takes the next one.
//
// Refinement variant: per system, FP64 residual r = b - A x, inner FP32 CG
// (FP32 values of the same layout, FP32 vectors, FP64 scalars and
// reductions) on r/|r| to the inner tolerance, FP64 update x += |r| d, until
// |r|/|b| <= rtol_outer.  Residual-replacement variant (symmetrically scaled
// FP32 matrix only): per system one FP32 CG whose recursive residual is
// replaced by the true FP64 one after each segment, the search direction
// kept (the per-system form of fsx_ir_cg_rr_solve).
//
// Row mapping inside a workgroup (chosen by the autotuner, fsx_batch_tune):
//   0: sliced rows -- one row per lane on the matrix's sliced layout
//      (coalesced, rows summed serially: bit-identical SpMV);
//   1: L lanes per row on the CSR arrays (L = 2, 4, 8), shuffle-combined;
//   2: one row per lane on the slice-offset layout (per-slice column
//      offsets and lane masks; serial row sums, bit-identical SpMV).
#include "fsx_spmv_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <vector>

struct fsx_batch {
	fsx_ctx * ctx = nullptr;
	const fsx_csr * A = nullptr;
	int64_t nsys = 0;
	std::vector<int64_t> h_rows;
	std::unique_ptr<fsx::dbuf> rows, order, queue;
	// execution configuration
	int wg = 256;
	int mapping = 0; // 0 sliced rows, 1 lanes per row, 2 slice offsets
	int lanes = 1;
	bool tuned = false;
	double tuned_ms = 0;
};

namespace fsx {

constexpr int kBatchMaxThreads = 1024;

struct batch_args {
	int64_t nsys;
	const int64_t * rows;
	const int * order;
	unsigned * queue;
	// matrix (FP64 and, for refinement, FP32 values of the same layout)
	const int32_t * rp;
	const int32_t * ci;
	const double * v64;
	const float * v32;
	const int32_t * sl_ptr;
	const int32_t * sl_col;
	const double * sl_v64;
	const float * sl_v32;
	const int64_t * dg_ptr;
	const int32_t * dg_off;
	const uint64_t * dg_mask;
	const double * dg_v64;
	const float * dg_v32;
	int mapping, lanes;
	// vectors (concatenated)
	const double * b;
	double * x;
	const double * dinv64;
	const float * dinv32;
	const double * sym; // refinement on S A S, S = D^{-1/2} (null: unscaled A32)
	double * r, * z, * p, * w;          // FP64 work
	float * fr, * fz, * fp, * fw, * fd; // FP32 inner work
	float * fb;                         // FP32 inner right-hand side
	// settings
	int maxiter;
	double rtol; // (double)(float) rtol
	int use_zero_guess;
	// refinement
	double rtol_outer;
	int max_outer;
	// residual replacement instead of restarts (a.sym set): the segment
	// reduction; max_outer bounds the segments
	int rr;
	double seg_rtol;
	// outputs
	int * iters;
	int * status;
	double * res;
	int * inner_iters;
};

// ---- workgroup helpers ------------------------------------------------------

// Lane 0's value of v, in scalar registers: a branch on it is uniform across
// the wave.  Every value that steers control flow in the persistent kernel
// goes through this (or is a kernel argument), so no loop below has a
// divergent exit or a branch the compiler could split into per-lane loops
// around the workgroup barriers.
__device__ __forceinline__ double lane0(double v) {
	const long long b = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
	const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
	return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ int lane0(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Sum of NV values over the workgroup in a fixed order; the result, uniform,
// in all threads.  Each wave's lanes all store lane 0's total (one value per
// slot, no divergent branch).
template<int NV>
__device__ __forceinline__ void wg_sum(double (&v)[NV], double * lds) {
	const int nw = blockDim.x >> 6;
	const int wv = threadIdx.x >> 6;
#pragma unroll
	for (int i = 0; i < NV; ++i) lds[i * 16 + wv] = lane0(wave_sum(v[i]));
	__syncthreads();
#pragma unroll
	for (int i = 0; i < NV; ++i) {
		double t = 0.0;
		for (int q = 0; q < nw; ++q) t += lds[i * 16 + q];
		v[i] = lane0(t);
	}
	__syncthreads();
}

// y = A x over rows [r0, r0 + n) of the concatenation; calls f(local row,
// row sum) once per row on exactly one thread.  `xl` points at the system's
// first entry; columns are global ids, so gathers index xl - r0.  The slice
// mappings use branch-free tiles of 8 entries (all loads issued before the
// first use; absent entries re-read the row's own x and are dropped by a
// select, bitwise the same as skipping them).
constexpr int kBatchTile = 4;
static_assert(kBatchTile <= kSliceUnroll, "offset/mask padding covers one tile");

template<class V, class X, class F>
__device__ __forceinline__ void wg_spmv(const batch_args & a, const V * sl_v,
                                        const V * v, const V * dg_v, int64_t r0,
                                        int64_t n, const X * xl, F && f) {
	const X * x = xl - r0;
	if (a.mapping == 2) {
		for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
			const int64_t r = r0 + i;
			const int64_t s = r / kSliceRows;
			const int lane = (int)(r & (kSliceRows - 1));
			const int64_t base = a.dg_ptr[s];
			const int wd = (int)((a.dg_ptr[s + 1] - base) / kSliceRows);
			const int64_t q0 = base / kSliceRows;
			double sum = 0.0;
			for (int j0 = 0; j0 < wd; j0 += kBatchTile) {
				int32_t off[kBatchTile];
				bool pr[kBatchTile];
				V vv[kBatchTile];
				X xv[kBatchTile];
#pragma unroll
				for (int k = 0; k < kBatchTile; ++k) {
					const int jj = min(j0 + k, wd - 1);
					// consecutive unconditional loads (the offset and mask arrays
					// are padded): wide scalar loads behind one wait
					off[k] = a.dg_off[q0 + j0 + k];
					const uint64_t m = a.dg_mask[q0 + j0 + k];
					pr[k] = (j0 + k < wd) & (((m >> lane) & 1) != 0);
					vv[k] = dg_v[base + (int64_t)jj * kSliceRows + lane];
				}
#pragma unroll
				for (int k = 0; k < kBatchTile; ++k) xv[k] = x[pr[k] ? r + off[k] : r];
#pragma unroll
				for (int k = 0; k < kBatchTile; ++k) {
					const double t = sum + (double)vv[k] * (double)xv[k];
					sum = pr[k] ? t : sum;
				}
			}
			f(i, sum);
		}
	}
	else if (a.mapping == 0) {
		for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
			const int64_t r = r0 + i;
			const int64_t s = r / kSliceRows;
			const int lane = (int)(r & (kSliceRows - 1));
			const int32_t base = a.sl_ptr[s];
			const int wd = (a.sl_ptr[s + 1] - base) / kSliceRows;
			double sum = 0.0;
			for (int j0 = 0; j0 < wd; j0 += kBatchTile) {
				int32_t c[kBatchTile];
				V vv[kBatchTile];
				X xv[kBatchTile];
#pragma unroll
				for (int k = 0; k < kBatchTile; ++k) {
					const int jj = min(j0 + k, wd - 1);
					c[k] = j0 + k < wd ? a.sl_col[base + jj * kSliceRows + lane] : -1;
					vv[k] = sl_v[base + jj * kSliceRows + lane];
				}
#pragma unroll
				for (int k = 0; k < kBatchTile; ++k) xv[k] = x[c[k] >= 0 ? (int64_t)c[k] : r];
#pragma unroll
				for (int k = 0; k < kBatchTile; ++k) {
					const double t = sum + (double)vv[k] * (double)xv[k];
					sum = c[k] >= 0 ? t : sum;
				}
			}
			f(i, sum);
		}
	}
	else {
		const int L = a.lanes;
		const int sub = threadIdx.x & (L - 1);
		const int groups = blockDim.x / L;
		// every lane of a group runs the same trip count (shuffles)
		for (int64_t i0 = 0; i0 < n; i0 += groups) {
			const int64_t i = i0 + threadIdx.x / L;
			double sum = 0.0;
			if (i < n) {
				const int64_t r = r0 + i;
				for (int32_t e = a.rp[r] + sub; e < a.rp[r + 1]; e += L)
					sum += (double)v[e] * (double)x[a.ci[e]];
			}
			for (int o = L / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, L);
			if (i < n && sub == 0) f(i, sum);
		}
	}
}

// ---- one CG solve of one system (cg.hh:43-139) --------------------------------
//
// T is the vector/storage type (double, or float for the refinement's inner
// solve), V the matrix value type; scalars and reductions are FP64.  The
// vector passes handle kBatchRows rows per thread per trip with every load
// issued first.  x += alpha p (cg.hh:108) runs in the p pass, which reads p
// anyway (80 instead of 88 bytes of vectors per row and iteration); the
// iteration that ends the loop applies it in a last pass.
constexpr int kBatchRows = 4;

template<bool JAC, class T, class V>
__device__ void wg_cg(const batch_args & a, const V * sl_v, const V * v,
                      const V * dg_v, int64_t r0, int64_t n, const T * b, T * x, T * r, T * z,
                      T * p, T * w, const T * dinv, int maxiter, double rtol,
                      bool zero_guess, double * lds, int & iters_out,
                      int & status_out, double & res_out) {
	const int tid = threadIdx.x, nt = blockDim.x;
	constexpr bool F64 = std::is_same_v<T, double>;
	double acc[2];
	double one[1];
	// b_norm (cg.hh:55-60)
	one[0] = 0.0;
	for (int64_t i = tid; i < n; i += nt) one[0] += (double)b[i] * (double)b[i];
	wg_sum<1>(one, lds);
	double bnorm = sqrt(one[0]);
	if (bnorm == 0.0) bnorm = 1.0;
	const double tol = rtol * bnorm;
	// initial residual (cg.hh:65-73)
	if (zero_guess) {
		for (int64_t i = tid; i < n; i += nt) {
			x[i] = T(0);
			r[i] = b[i];
		}
	}
	else {
		wg_spmv(a, sl_v, v, dg_v, r0, n, (const T *)x, [&](int64_t i, double s) {
			r[i] = b[i] - (T)s;
		});
	}
	__syncthreads();
	one[0] = 0.0;
	for (int64_t i = tid; i < n; i += nt) one[0] += (double)r[i] * (double)r[i];
	wg_sum<1>(one, lds);
	double res = sqrt(one[0]);
	int iters = 0, status = FSX_UNKNOWN;
	if (res < tol) {
		iters_out = 0;
		status_out = FSX_CONVERGED_RTOL;
		res_out = res;
		return;
	}
	// z = P r, rho = z.r, p = z (cg.hh:84-90)
	const T * zz = JAC ? z : r;
	one[0] = 0.0;
	for (int64_t i = tid; i < n; i += nt) {
		T zi = r[i];
		if constexpr (JAC) {
			if constexpr (F64) zi = 0.0 + dinv[i] * r[i];
			else zi = (T)((double)dinv[i] * (double)r[i]);
			z[i] = zi;
		}
		p[i] = zi;
		one[0] += (double)zi * (double)r[i];
	}
	wg_sum<1>(one, lds);
	double rho = one[0], rho0, alpha = 0.0;
	bool ran = false;
	for (int it = 0; it < maxiter; ++it) {
		ran = true;
		// w = A p, alpha = rho / (w.p)  (cg.hh:95-106)
		one[0] = 0.0;
		wg_spmv(a, sl_v, v, dg_v, r0, n, (const T *)p, [&](int64_t i, double s) {
			const T wi = (T)s;
			w[i] = wi;
			one[0] += (double)wi * (double)p[i];
		});
		wg_sum<1>(one, lds);
		alpha = rho / one[0];
		// r -= alpha w; |r|; z = P r; r.z  (cg.hh:109-127)
		acc[0] = 0.0;
		acc[1] = 0.0;
		for (int64_t i0 = tid; i0 < n; i0 += (int64_t)kBatchRows * nt) {
			T wv[kBatchRows], rv[kBatchRows], dv[kBatchRows];
#pragma unroll
			for (int k = 0; k < kBatchRows; ++k) {
				const int64_t i = i0 + (int64_t)k * nt;
				if (i < n) {
					wv[k] = w[i];
					rv[k] = r[i];
					if constexpr (JAC) dv[k] = dinv[i];
				}
			}
#pragma unroll
			for (int k = 0; k < kBatchRows; ++k) {
				const int64_t i = i0 + (int64_t)k * nt;
				if (i < n) {
					T ri;
					if constexpr (F64) ri = (-alpha) * wv[k] + rv[k];
					else ri = (T)((-alpha) * (double)wv[k] + (double)rv[k]);
					r[i] = ri;
					acc[0] += (double)ri * (double)ri;
					if constexpr (JAC) {
						T zi;
						if constexpr (F64) zi = 0.0 + dv[k] * ri;
						else zi = (T)((double)dv[k] * (double)ri);
						z[i] = zi;
						acc[1] += (double)ri * (double)zi;
					}
				}
			}
		}
		wg_sum<2>(acc, lds);
		res = sqrt(acc[0]);
		if (res < tol) {
			iters = it + 1;
			status = FSX_CONVERGED_RTOL;
			break;
		}
		if (it + 1 >= maxiter) break;
		rho0 = rho;
		rho = JAC ? acc[1] : acc[0];
		const double beta = rho / rho0;
		// x += alpha p; p = beta p + z  (cg.hh:108, 130)
		for (int64_t i0 = tid; i0 < n; i0 += (int64_t)kBatchRows * nt) {
			T pv[kBatchRows], zv[kBatchRows], xv[kBatchRows];
#pragma unroll
			for (int k = 0; k < kBatchRows; ++k) {
				const int64_t i = i0 + (int64_t)k * nt;
				if (i < n) {
					pv[k] = p[i];
					zv[k] = zz[i];
					xv[k] = x[i];
				}
			}
#pragma unroll
			for (int k = 0; k < kBatchRows; ++k) {
				const int64_t i = i0 + (int64_t)k * nt;
				if (i < n) {
					if constexpr (F64) {
						x[i] = alpha * pv[k] + xv[k];
						p[i] = beta * pv[k] + zv[k];
					}
					else {
						x[i] = (T)(alpha * (double)pv[k] + (double)xv[k]);
						p[i] = (T)(beta * (double)pv[k] + (double)zv[k]);
					}
				}
			}
		}
		__syncthreads();
	}
	// the last iteration's x += alpha p
	if (ran) {
		for (int64_t i = tid; i < n; i += nt) {
			if constexpr (F64) x[i] = alpha * p[i] + x[i];
			else x[i] = (T)(alpha * (double)p[i] + (double)x[i]);
		}
		__syncthreads();
	}
	if (iters == 0) status = FSX_DIVERGED_ITERS;
	iters_out = iters;
	status_out = status;
	res_out = res;
}

// One segment of the unpreconditioned CG with residual replacement: from
// r (recursive residual), p (search direction), rho = |r|^2, updates
// r -= alpha w, then y += alpha p and p = beta p + r (beta = |r_new|^2 /
// |r|^2, the two-reduction form), until |r| < tol or maxiter updates.  On
// return r, p, y and rho hold the state the replacement continues from.
template<class T, class V>
__device__ int wg_cg_segment(const batch_args & a, const V * sl_v, const V * v,
                             const V * dg_v, int64_t r0, int64_t n, T * y, T * r, T * p,
                             T * w, double & rho, double tol, int maxiter, double * lds) {
	const int tid = threadIdx.x, nt = blockDim.x;
	double one[1];
	int it = 0;
	for (;;) {
		one[0] = 0.0;
		wg_spmv(a, sl_v, v, dg_v, r0, n, (const T *)p, [&](int64_t i, double s) {
			const T wi = (T)s;
			w[i] = wi;
			one[0] += (double)wi * (double)p[i];
		});
		wg_sum<1>(one, lds);
		const double alpha = rho / one[0];
		one[0] = 0.0;
		for (int64_t i0 = tid; i0 < n; i0 += (int64_t)kBatchRows * nt) {
			T wv[kBatchRows], rv[kBatchRows];
#pragma unroll
			for (int k = 0; k < kBatchRows; ++k) {
				const int64_t i = i0 + (int64_t)k * nt;
				if (i < n) {
					wv[k] = w[i];
					rv[k] = r[i];
				}
			}
#pragma unroll
			for (int k = 0; k < kBatchRows; ++k) {
				const int64_t i = i0 + (int64_t)k * nt;
				if (i < n) {
					const T ri = (T)((-alpha) * (double)wv[k] + (double)rv[k]);
					r[i] = ri;
					one[0] += (double)ri * (double)ri;
				}
			}
		}
		wg_sum<1>(one, lds);
		const double beta = one[0] / rho;
		rho = one[0];
		for (int64_t i0 = tid; i0 < n; i0 += (int64_t)kBatchRows * nt) {
			T pv[kBatchRows], rv[kBatchRows], yv[kBatchRows];
#pragma unroll
			for (int k = 0; k < kBatchRows; ++k) {
				const int64_t i = i0 + (int64_t)k * nt;
				if (i < n) {
					pv[k] = p[i];
					rv[k] = r[i];
					yv[k] = y[i];
				}
			}
#pragma unroll
			for (int k = 0; k < kBatchRows; ++k) {
				const int64_t i = i0 + (int64_t)k * nt;
				if (i < n) {
					y[i] = (T)(alpha * (double)pv[k] + (double)yv[k]);
					p[i] = (T)(beta * (double)pv[k] + (double)rv[k]);
				}
			}
		}
		__syncthreads();
		++it;
		if (sqrt(rho) < tol || it >= maxiter) break;
	}
	return it;
}

template<bool JAC, bool IR>
__global__ __launch_bounds__(kBatchMaxThreads) void batched_cg_kernel(batch_args a) {
	__shared__ double lds[2 * 16];
	__shared__ int s_q[16];
	const int wv = threadIdx.x >> 6;
	for (;;) {
		// Next queue slot.  Every thread takes part in the atomic (thread 0
		// adds 1, the rest 0) and each wave stores its lane 0's old value, so
		// no branch depends on the thread index; slot 0 holds thread 0's.
		const unsigned old = atomicAdd(a.queue, threadIdx.x == 0 ? 1u : 0u);
		s_q[wv] = lane0((int)old);
		__syncthreads();
		const int q = lane0(s_q[0]);
		__syncthreads();
		if (q >= a.nsys) break;
		const int sys = a.order[q];
		const int64_t r0 = a.rows[sys];
		const int64_t n = a.rows[sys + 1] - r0;
		int iters = 0, status = FSX_UNKNOWN, inner = 0;
		double res = 0.0;
		if constexpr (!IR) {
			wg_cg<JAC, double, double>(a, a.sl_v64, a.v64, a.dg_v64, r0, n, a.b + r0,
			                           a.x + r0,
			                           a.r + r0, a.z + r0, a.p + r0, a.w + r0,
			                           a.dinv64 + r0, a.maxiter, a.rtol,
			                           a.use_zero_guess != 0, lds, iters, status, res);
		}
		else if (a.rr) {
			// FP32 CG of this system with FP64 residual replacement
			const double * b = a.b + r0;
			double * x = a.x + r0;
			double * r = a.r + r0;
			const double * sc = a.sym + r0;
			float * rh = a.fr + r0;
			float * p = a.fp + r0;
			float * y = a.fd + r0;
			double acc[1] = {0.0};
			for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc[0] += b[i] * b[i];
			wg_sum<1>(acc, lds);
			double bnorm = sqrt(acc[0]);
			if (bnorm == 0.0) bnorm = 1.0;
			// r = b - A64 x and |r|
			auto residual = [&]() {
				acc[0] = 0.0;
				wg_spmv(a, a.sl_v64, a.v64, a.dg_v64, r0, n, (const double *)x,
				        [&](int64_t i, double s) {
					const double ri = b[i] - s;
					r[i] = ri;
					acc[0] += ri * ri;
				});
				wg_sum<1>(acc, lds);
				return sqrt(acc[0]);
			};
			res = residual();
			int seg = 0;
			if (res / bnorm <= a.rtol_outer || res == 0.0)
				status = FSX_CONVERGED_RTOL;
			else {
				const double c = res, inv = 1.0 / res;
				// rh = (float)(S r / c), p = rh, y = 0
				acc[0] = 0.0;
				for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
					const float vi = (float)(sc[i] * (r[i] * inv));
					rh[i] = vi;
					p[i] = vi;
					y[i] = 0.0f;
					acc[0] += (double)vi * (double)vi;
				}
				wg_sum<1>(acc, lds);
				double rho = acc[0];
				for (;;) {
					if (seg >= a.max_outer) {
						status = FSX_DIVERGED_ITERS;
						break;
					}
					if (!(rho > 0.0)) {
						status = FSX_DIVERGED_BREAKDOWN;
						break;
					}
					const double rh0 = sqrt(rho);
					const double tol = fmax(a.seg_rtol * rh0,
					                        0.5 * a.rtol_outer * bnorm * rh0 / res);
					inner += wg_cg_segment<float, float>(a, a.sl_v32, a.v32, a.dg_v32, r0, n,
					                                     y, rh, p, a.fw + r0, rho, tol,
					                                     a.maxiter, lds);
					++seg;
					for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
						x[i] = x[i] + c * (sc[i] * (double)y[i]);
					__syncthreads();
					res = residual();
					if (res / bnorm <= a.rtol_outer || res == 0.0) {
						status = FSX_CONVERGED_RTOL;
						break;
					}
					// the replacement: rh <- (float)(S r / c), p += rh_new - rh_old
					acc[0] = 0.0;
					for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
						const float vi = (float)(sc[i] * (r[i] * inv));
						p[i] = (float)((double)p[i] + ((double)vi - (double)rh[i]));
						rh[i] = vi;
						y[i] = 0.0f;
						acc[0] += (double)vi * (double)vi;
					}
					wg_sum<1>(acc, lds);
					rho = acc[0];
				}
			}
			iters = seg;
		}
		else {
			// FP32-inner / FP64-outer refinement of this system
			const double * b = a.b + r0;
			double * x = a.x + r0;
			double * r = a.r + r0;
			double acc[1] = {0.0};
			for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc[0] += b[i] * b[i];
			wg_sum<1>(acc, lds);
			double bnorm = sqrt(acc[0]);
			if (bnorm == 0.0) bnorm = 1.0;
			int outer = 0;
			for (;;) {
				// r = b - A64 x, |r|
				acc[0] = 0.0;
				wg_spmv(a, a.sl_v64, a.v64, a.dg_v64, r0, n, (const double *)x,
				        [&](int64_t i, double s) {
					const double ri = b[i] - s;
					r[i] = ri;
					acc[0] += ri * ri;
				});
				wg_sum<1>(acc, lds);
				res = sqrt(acc[0]);
				if (res / bnorm <= a.rtol_outer || res == 0.0) {
					status = FSX_CONVERGED_RTOL;
					break;
				}
				if (outer >= a.max_outer) {
					status = FSX_DIVERGED_ITERS;
					break;
				}
				const double inv = 1.0 / res;
				if (a.sym)
					for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
						a.fb[r0 + i] = (float)(a.sym[r0 + i] * (r[i] * inv));
				else
					for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
						a.fb[r0 + i] = (float)(r[i] * inv);
				__syncthreads();
				int it_in = 0, st_in = 0;
				double res_in = 0.0;
				wg_cg<JAC, float, float>(a, a.sl_v32, a.v32, a.dg_v32, r0, n, a.fb + r0,
				                         a.fd + r0,
				                         a.fr + r0, a.fz + r0, a.fp + r0, a.fw + r0,
				                         a.dinv32 + r0, a.maxiter, a.rtol, true, lds,
				                         it_in, st_in, res_in);
				inner += it_in ? it_in : a.maxiter;
				if (a.sym)
					for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
						x[i] = x[i] + res * (a.sym[r0 + i] * (double)a.fd[r0 + i]);
				else
					for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
						x[i] = x[i] + res * (double)a.fd[r0 + i];
				__syncthreads();
				++outer;
			}
			iters = outer;
		}
		// all lanes of wave 0 store the same (uniform) values
		if (lane0(wv) == 0) {
			a.iters[sys] = iters;
			a.status[sys] = status;
			a.res[sys] = res;
			if (a.inner_iters) a.inner_iters[sys] = inner;
		}
	}
}

// Device check that every row of system s only references columns of s.
__global__ void batch_check_kernel(int64_t nsys, const int64_t * rows,
                                   const int32_t * rp, const int32_t * ci,
                                   unsigned * bad) {
	const int64_t s = blockIdx.x;
	if (s >= nsys) return;
	const int64_t lo = rows[s], hi = rows[s + 1];
	for (int64_t r = lo + threadIdx.x; r < hi; r += blockDim.x)
		for (int32_t e = rp[r]; e < rp[r + 1]; ++e)
			if (ci[e] < lo || ci[e] >= hi) atomicOr(bad, 1u);
}

} // namespace fsx

using namespace fsx;

namespace {

template<bool JAC, bool IR>
int occupancy(int wg) {
	int nb = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
			&nb, reinterpret_cast<const void *>(&batched_cg_kernel<JAC, IR>), wg, 0) !=
	    hipSuccess)
		return 1;
	return std::max(1, nb);
}

template<bool JAC, bool IR>
void launch(fsx_ctx * ctx, fsx_batch * B, const batch_args & a) {
	const int per_cu = occupancy<JAC, IR>(B->wg);
	const int64_t grid = std::min<int64_t>((int64_t)ctx->num_cus * per_cu, B->nsys);
	FSX_HIP(hipMemsetAsync(B->queue->p, 0, sizeof(unsigned), ctx->stream));
	hipLaunchKernelGGL((batched_cg_kernel<JAC, IR>), dim3((unsigned)std::max<int64_t>(1, grid)),
	                   dim3(B->wg), 0, ctx->stream, a);
	FSX_HIP(hipGetLastError());
	ctx->launches++;
}

void check_batch(const fsx_batch * B) {
	FSX_REQUIRE(B && B->A, "null batch");
}

bool has_sliced(const fsx_csr * A) { return A->idx->sl_ptr && A->sl_vals; }
bool has_sdia(const fsx_csr * A) { return A->idx->dg_ptr && A->dg_vals; }

struct batch_outputs {
	std::unique_ptr<dbuf> iters, status, res, inner;
	explicit batch_outputs(int64_t nsys) {
		iters = std::make_unique<dbuf>(std::max<int64_t>(1, nsys) * sizeof(int));
		status = std::make_unique<dbuf>(std::max<int64_t>(1, nsys) * sizeof(int));
		res = std::make_unique<dbuf>(std::max<int64_t>(1, nsys) * sizeof(double));
		inner = std::make_unique<dbuf>(std::max<int64_t>(1, nsys) * sizeof(int));
	}
};

batch_args base_args(fsx_batch * B, const fsx_csr * A32) {
	const fsx_csr * A = B->A;
	ensure_dg_fresh(B->ctx, A);
	if (A32) ensure_dg_fresh(B->ctx, A32);
	batch_args a{};
	a.nsys = B->nsys;
	a.rows = (const int64_t *)B->rows->p;
	a.order = (const int *)B->order->p;
	a.queue = (unsigned *)B->queue->p;
	a.rp = (const int32_t *)A->rowptr();
	a.ci = (const int32_t *)A->colind();
	a.v64 = (const double *)A->vals->p;
	a.sl_ptr = A->idx->sl_ptr ? (const int32_t *)A->idx->sl_ptr->p : nullptr;
	a.sl_col = A->idx->sl_col ? (const int32_t *)A->idx->sl_col->p : nullptr;
	a.sl_v64 = A->sl_vals ? (const double *)A->sl_vals->p : nullptr;
	a.dg_ptr = A->idx->dg_ptr ? (const int64_t *)A->idx->dg_ptr->p : nullptr;
	a.dg_off = A->idx->dg_off ? (const int32_t *)A->idx->dg_off->p : nullptr;
	a.dg_mask = A->idx->dg_mask ? (const uint64_t *)A->idx->dg_mask->p : nullptr;
	a.dg_v64 = A->dg_vals ? (const double *)A->dg_vals->p : nullptr;
	if (A32) {
		a.v32 = (const float *)A32->vals->p;
		a.sl_v32 = A32->sl_vals ? (const float *)A32->sl_vals->p : nullptr;
		a.dg_v32 = A32->dg_vals ? (const float *)A32->dg_vals->p : nullptr;
	}
	a.mapping = B->mapping;
	a.lanes = B->lanes;
	return a;
}

// Work vectors pooled on the batch (stable addresses; sized to the total rows).
struct batch_work {
	std::unique_ptr<dbuf> r, z, p, w, fr, fz, fp, fw, fd, fb, dinv32, sym;
};
std::map<const fsx_batch *, batch_work> g_work;

batch_work & work_for(fsx_batch * B, bool ir) {
	auto & W = g_work[B];
	const int64_t N = B->A->nrows;
	auto mk = [&](std::unique_ptr<dbuf> & d, size_t el) {
		if (!d) d = std::make_unique<dbuf>(std::max<int64_t>(1, N) * el);
	};
	mk(W.r, 8);
	mk(W.z, 8);
	mk(W.p, 8);
	mk(W.w, 8);
	if (ir) {
		mk(W.fr, 4); mk(W.fz, 4); mk(W.fp, 4); mk(W.fw, 4); mk(W.fd, 4); mk(W.fb, 4);
	}
	return W;
}

void copy_out(fsx_ctx * ctx, const batch_outputs & o, int64_t nsys, int * iters,
              int * status, double * res, int * inner) {
	if (nsys == 0) return;
	if (iters)
		FSX_HIP(hipMemcpyAsync(iters, o.iters->p, nsys * sizeof(int),
		                       hipMemcpyDeviceToHost, ctx->stream));
	if (status)
		FSX_HIP(hipMemcpyAsync(status, o.status->p, nsys * sizeof(int),
		                       hipMemcpyDeviceToHost, ctx->stream));
	if (res)
		FSX_HIP(hipMemcpyAsync(res, o.res->p, nsys * sizeof(double),
		                       hipMemcpyDeviceToHost, ctx->stream));
	if (inner)
		FSX_HIP(hipMemcpyAsync(inner, o.inner->p, nsys * sizeof(int),
		                       hipMemcpyDeviceToHost, ctx->stream));
	FSX_HIP(hipStreamSynchronize(ctx->stream));
}

void set_order(fsx_ctx * ctx, fsx_batch * B, const std::vector<int> & ord) {
	FSX_HIP(hipMemcpyAsync(B->order->p, ord.data(), ord.size() * sizeof(int),
	                       hipMemcpyHostToDevice, ctx->stream));
	FSX_HIP(hipStreamSynchronize(ctx->stream));
}

std::vector<int> lpt_order(const fsx_batch * B) {
	std::vector<int> ord(B->nsys);
	std::iota(ord.begin(), ord.end(), 0);
	std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
		return B->h_rows[a + 1] - B->h_rows[a] > B->h_rows[b + 1] - B->h_rows[b];
	});
	return ord;
}

} // namespace

extern "C" {

int fsx_batch_create(fsx_ctx * ctx, const fsx_csr * A, int64_t nsys,
                     const int64_t * sys_rows, fsx_batch ** out) {
	return guard([&] {
		FSX_REQUIRE(ctx && A && sys_rows && out, "null argument");
		FSX_REQUIRE(nsys >= 0, "negative system count");
		FSX_REQUIRE(A->it == FSX_I32 && A->vt == FSX_F64,
		            "batched solves take an int32-indexed FP64 matrix");
		FSX_REQUIRE(A->nrows == A->ncols, "batch matrix must be square");
		FSX_REQUIRE(sys_rows[0] == 0 && sys_rows[nsys] == A->nrows,
		            "system row offsets must start at 0 and end at nrows");
		for (int64_t s = 0; s < nsys; ++s)
			FSX_REQUIRE(sys_rows[s + 1] >= sys_rows[s], "system row offsets not monotone");
		FSX_REQUIRE(nsys < ((int64_t)1 << 31), "too many systems");
		auto B = std::make_unique<fsx_batch>();
		B->ctx = ctx;
		B->A = A;
		B->nsys = nsys;
		B->h_rows.assign(sys_rows, sys_rows + nsys + 1);
		B->rows = std::make_unique<dbuf>((nsys + 1) * sizeof(int64_t));
		B->order = std::make_unique<dbuf>(std::max<int64_t>(1, nsys) * sizeof(int));
		B->queue = std::make_unique<dbuf>(256);
		FSX_HIP(hipMemcpyAsync(B->rows->p, sys_rows, (nsys + 1) * sizeof(int64_t),
		                       hipMemcpyHostToDevice, ctx->stream));
		// systems must be independent: no column outside the system's rows
		FSX_HIP(hipMemsetAsync(B->queue->p, 0, sizeof(unsigned), ctx->stream));
		if (nsys > 0)
			hipLaunchKernelGGL(batch_check_kernel, dim3((unsigned)nsys), dim3(256), 0,
			                   ctx->stream, nsys, (const int64_t *)B->rows->p,
			                   (const int32_t *)A->rowptr(), (const int32_t *)A->colind(),
			                   (unsigned *)B->queue->p);
		FSX_HIP(hipGetLastError());
		unsigned bad = 0;
		FSX_HIP(hipMemcpyAsync(&bad, B->queue->p, sizeof(unsigned), hipMemcpyDeviceToHost,
		                       ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		FSX_REQUIRE(bad == 0, "a system's rows reference columns of another system");
		set_order(ctx, B.get(), lpt_order(B.get()));
		B->mapping = has_sdia(A) ? 2 : has_sliced(A) ? 0 : 1;
		B->lanes = B->mapping == 1 ? 4 : 1;
		ctx_retain(ctx);
		*out = B.release();
	});
}

int fsx_batch_destroy(fsx_batch * B) {
	return guard([&] {
		if (!B) return;
		fsx_ctx * c = B->ctx;
		(void)hipStreamSynchronize(c->stream);
		g_work.erase(B);
		delete B;
		ctx_release(c);
	});
}

int fsx_batch_config(const fsx_batch * B, int * wg, int * mapping, int * lanes) {
	return guard([&] {
		check_batch(B);
		if (wg) *wg = B->wg;
		if (mapping) *mapping = B->mapping;
		if (lanes) *lanes = B->lanes;
	});
}

int fsx_batch_set_config(fsx_batch * B, int wg, int mapping, int lanes) {
	return guard([&] {
		check_batch(B);
		FSX_REQUIRE(wg == 64 || wg == 128 || wg == 256 || wg == 512 || wg == 1024,
		            "workgroup size must be 64, 128, 256, 512 or 1024");
		FSX_REQUIRE(mapping >= 0 && mapping <= 2, "mapping must be 0, 1 or 2");
		if (mapping == 0)
			FSX_REQUIRE(has_sliced(B->A), "sliced mapping needs the matrix's sliced layout");
		else if (mapping == 2)
			FSX_REQUIRE(has_sdia(B->A),
			            "slice-offset mapping needs the matrix's slice-offset layout");
		else
			FSX_REQUIRE(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 ||
			                lanes == 16,
			            "lanes must be 1, 2, 4, 8 or 16");
		B->wg = wg;
		B->mapping = mapping;
		B->lanes = mapping == 1 ? lanes : 1;
	});
}

// FP64 batched CG of the systems listed in B's order, using W's work arrays.
static void batched_cg_run(fsx_ctx * ctx, fsx_batch * B, batch_work & W,
                           const fsx_vec * dinv, const fsx_vec * b, fsx_vec * x,
                           const fsx_settings * s, int * iters, int * status,
                           double * res) {
		batch_outputs o(B->nsys);
		batch_args a = base_args(B, nullptr);
		a.b = (const double *)b->ptr();
		a.x = (double *)x->ptr();
		a.dinv64 = dinv ? (const double *)dinv->ptr() : nullptr;
		a.r = (double *)W.r->p;
		a.z = (double *)W.z->p;
		a.p = (double *)W.p->p;
		a.w = (double *)W.w->p;
		a.maxiter = s->maxiter;
		a.rtol = (double)s->rtol;
		a.use_zero_guess = s->use_zero_guess;
		a.iters = (int *)o.iters->p;
		a.status = (int *)o.status->p;
		a.res = (double *)o.res->p;
		a.inner_iters = nullptr;
		if (B->nsys > 0) {
			if (dinv) launch<true, false>(ctx, B, a);
			else launch<false, false>(ctx, B, a);
		}
		copy_out(ctx, o, B->nsys, iters, status, res, nullptr);
}

int fsx_batched_cg(fsx_ctx * ctx, fsx_batch * B, const fsx_vec * dinv,
                   const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                   int * iters, int * status, double * res) {
	return guard([&] {
		FSX_REQUIRE(ctx && s, "null argument");
		check_batch(B);
		check_vec(b, "b");
		check_vec(x, "x");
		check_same(b, x);
		FSX_REQUIRE((int64_t)b->n == B->A->nrows && b->dt == FSX_F64,
		            "b and x: FP64, one entry per batch row");
		FSX_REQUIRE(b->ptr() != x->ptr() || b->n == 0, "b and x must differ");
		if (dinv) check_same(dinv, b);
		FSX_REQUIRE(s->maxiter >= 0, "maxiter must be >= 0");
		batched_cg_run(ctx, B, work_for(B, false), dinv, b, x, s, iters, status, res);
	});
}

int fsx_batched_ir_cg(fsx_ctx * ctx, fsx_batch * B, const fsx_csr * A32,
                      int jacobi, const fsx_vec * b, fsx_vec * x,
                      double rtol_outer, int max_outer, const fsx_settings * inner,
                      int * outer_iters, int * inner_iters, int * status,
                      double * res) {
	return guard([&] {
		FSX_REQUIRE(ctx && A32 && inner, "null argument");
		check_batch(B);
		check_vec(b, "b");
		check_vec(x, "x");
		check_same(b, x);
		FSX_REQUIRE((int64_t)b->n == B->A->nrows && b->dt == FSX_F64,
		            "b and x: FP64, one entry per batch row");
		FSX_REQUIRE(A32->vt == FSX_F32 && A32->idx == B->A->idx,
		            "A32 must be fsx_csr_convert(A, FSX_F32) of the batch matrix");
		FSX_REQUIRE(B->mapping != 0 || A32->sl_vals, "A32 lacks the sliced layout");
		FSX_REQUIRE(B->mapping != 2 || A32->dg_vals, "A32 lacks the slice-offset layout");
		FSX_REQUIRE(jacobi >= 0 && jacobi <= 2, "jacobi must be 0, 1 or 2");
		auto & W = work_for(B, true);
		if (jacobi == 2) {
			// S = D^{-1/2} of the FP64 batch matrix (A32 holds S A S, written by
			// fsx_csr_set_values_f32_sym_scaled)
			if (!W.sym) W.sym = std::make_unique<dbuf>(std::max<int64_t>(1, B->A->nrows) * 8);
			dbuf flag(sizeof(int));
			FSX_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), ctx->stream));
			sym_scale_vector(ctx, B->A, (double *)W.sym->p, (int *)flag.p);
			int bad = 0;
			FSX_HIP(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
			FSX_HIP(hipStreamSynchronize(ctx->stream));
			ctx->host_syncs++;
			FSX_REQUIRE(!bad, "a diagonal entry of the batch matrix is missing or not positive");
		}
		if (jacobi == 1 && !W.dinv32) {
			W.dinv32 = std::make_unique<dbuf>(std::max<int64_t>(1, B->A->nrows) * 4);
			fsx_vec * tmp = nullptr;
			int rc = fsx_vec_wrap(ctx, W.dinv32->p, B->A->nrows, FSX_F32, &tmp);
			if (rc) throw error(rc, get_error());
			rc = fsx_csr_diag_inv(ctx, B->A, tmp);
			fsx_vec_release(tmp);
			if (rc) throw error(rc, get_error());
		}
		batch_outputs o(B->nsys);
		batch_args a = base_args(B, A32);
		a.b = (const double *)b->ptr();
		a.x = (double *)x->ptr();
		a.dinv32 = jacobi == 1 ? (const float *)W.dinv32->p : nullptr;
		a.sym = jacobi == 2 ? (const double *)W.sym->p : nullptr;
		a.r = (double *)W.r->p;
		a.fr = (float *)W.fr->p;
		a.fz = (float *)W.fz->p;
		a.fp = (float *)W.fp->p;
		a.fw = (float *)W.fw->p;
		a.fd = (float *)W.fd->p;
		a.fb = (float *)W.fb->p;
		a.maxiter = inner->maxiter;
		a.rtol = (double)inner->rtol;
		a.use_zero_guess = 1;
		a.rtol_outer = rtol_outer;
		a.max_outer = max_outer;
		a.iters = (int *)o.iters->p;
		a.status = (int *)o.status->p;
		a.res = (double *)o.res->p;
		a.inner_iters = (int *)o.inner->p;
		if (B->nsys > 0) {
			if (jacobi == 1) launch<true, true>(ctx, B, a);
			else launch<false, true>(ctx, B, a);
		}
		copy_out(ctx, o, B->nsys, outer_iters, status, res, inner_iters);
	});
}

int fsx_batched_rr_cg(fsx_ctx * ctx, fsx_batch * B, const fsx_csr * A32,
                      const fsx_vec * b, fsx_vec * x, double rtol, int max_segments,
                      const fsx_settings * inner, int * segments, int * inner_iters,
                      int * status, double * res) {
	return guard([&] {
		FSX_REQUIRE(ctx && A32 && inner, "null argument");
		check_batch(B);
		check_vec(b, "b");
		check_vec(x, "x");
		check_same(b, x);
		FSX_REQUIRE((int64_t)b->n == B->A->nrows && b->dt == FSX_F64,
		            "b and x: FP64, one entry per batch row");
		FSX_REQUIRE(b->ptr() != x->ptr() || b->n == 0, "b and x must differ");
		FSX_REQUIRE(A32->vt == FSX_F32 && A32->idx == B->A->idx,
		            "A32 must be fsx_csr_convert(A, FSX_F32) of the batch matrix holding "
		            "fsx_csr_set_values_f32_sym_scaled values");
		FSX_REQUIRE(B->mapping != 0 || A32->sl_vals, "A32 lacks the sliced layout");
		FSX_REQUIRE(B->mapping != 2 || A32->dg_vals, "A32 lacks the slice-offset layout");
		FSX_REQUIRE(inner->rtol > 0.0 && inner->rtol < 1.0,
		            "the segment reduction inner->rtol must lie in (0, 1)");
		FSX_REQUIRE(inner->maxiter >= 1, "inner->maxiter must be at least 1");
		FSX_REQUIRE(max_segments >= 0, "max_segments must not be negative");
		auto & W = work_for(B, true);
		// S = D^{-1/2} of the FP64 batch matrix
		if (!W.sym) W.sym = std::make_unique<dbuf>(std::max<int64_t>(1, B->A->nrows) * 8);
		dbuf flag(sizeof(int));
		FSX_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), ctx->stream));
		sym_scale_vector(ctx, B->A, (double *)W.sym->p, (int *)flag.p);
		int bad = 0;
		FSX_HIP(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		ctx->host_syncs++;
		FSX_REQUIRE(!bad, "a diagonal entry of the batch matrix is missing or not positive");
		batch_outputs o(B->nsys);
		batch_args a = base_args(B, A32);
		a.b = (const double *)b->ptr();
		a.x = (double *)x->ptr();
		a.sym = (const double *)W.sym->p;
		a.r = (double *)W.r->p;
		a.fr = (float *)W.fr->p;
		a.fz = (float *)W.fz->p;
		a.fp = (float *)W.fp->p;
		a.fw = (float *)W.fw->p;
		a.fd = (float *)W.fd->p;
		a.fb = (float *)W.fb->p;
		a.maxiter = inner->maxiter;
		a.rtol = (double)inner->rtol;
		a.use_zero_guess = 0;
		a.rtol_outer = rtol;
		a.max_outer = max_segments;
		a.rr = 1;
		a.seg_rtol = (double)inner->rtol;
		a.iters = (int *)o.iters->p;
		a.status = (int *)o.status->p;
		a.res = (double *)o.res->p;
		a.inner_iters = (int *)o.inner->p;
		if (B->nsys > 0) launch<false, true>(ctx, B, a);
		copy_out(ctx, o, B->nsys, segments, status, res, inner_iters);
	});
}

// Autotuner: times each candidate configuration (workgroup size x row
// mapping) on a sample of the batch's systems for a fixed amount of work
// (`iters` CG iterations each, tolerance 0 so none stops early) and keeps the
// fastest.  Results are cached per sparsity profile (mean nonzeros per row,
// log2 of the median system size).
int fsx_batch_tune(fsx_ctx * ctx, fsx_batch * B, const fsx_vec * dinv,
                   const fsx_vec * b, fsx_vec * x, int sample, int iters,
                   double * best_ms) {
	return guard([&] {
		FSX_REQUIRE(ctx, "null context");
		check_batch(B);
		FSX_REQUIRE(sample > 0 && iters > 0, "sample and iters must be > 0");
		check_vec(b, "b");
		check_vec(x, "x");
		check_same(b, x);
		FSX_REQUIRE((int64_t)b->n == B->A->nrows && b->dt == FSX_F64,
		            "b and x: FP64, one entry per batch row");
		FSX_REQUIRE(b->ptr() != x->ptr(), "b and x must differ");
		if (dinv) check_same(dinv, b);
		if (B->nsys == 0) return;
		static std::map<std::pair<int, int>, std::tuple<int, int, int, double>> cache;
		std::vector<int64_t> sizes(B->nsys);
		for (int64_t s = 0; s < B->nsys; ++s) sizes[s] = B->h_rows[s + 1] - B->h_rows[s];
		std::vector<int64_t> sorted = sizes;
		std::nth_element(sorted.begin(), sorted.begin() + sorted.size() / 2, sorted.end());
		const int64_t med = std::max<int64_t>(1, sorted[sorted.size() / 2]);
		const int nnz_row = B->A->nrows ? (int)std::lround((double)B->A->nnz / B->A->nrows) : 0;
		const std::pair<int, int> key{nnz_row, (int)std::floor(std::log2((double)med))};
		auto hit = cache.find(key);
		if (hit != cache.end()) {
			std::tie(B->wg, B->mapping, B->lanes, B->tuned_ms) = hit->second;
			B->tuned = true;
			if (best_ms) *best_ms = B->tuned_ms;
			return;
		}
		// sample: systems spread over the size-sorted order
		std::vector<int> ord = lpt_order(B);
		const int64_t m = std::min<int64_t>(sample, B->nsys);
		std::vector<int> samp(m);
		for (int64_t i = 0; i < m; ++i) samp[i] = ord[(i * B->nsys) / m];
		fsx_batch S;
		S.ctx = ctx;
		S.A = B->A;
		S.nsys = m;
		S.h_rows = B->h_rows;
		// the kernel indexes rows[order[q]]: keep the full row table, sample ids
		S.rows = std::make_unique<dbuf>((B->nsys + 1) * sizeof(int64_t));
		S.order = std::make_unique<dbuf>(m * sizeof(int));
		S.queue = std::make_unique<dbuf>(256);
		FSX_HIP(hipMemcpyAsync(S.rows->p, B->rows->p, (B->nsys + 1) * sizeof(int64_t),
		                       hipMemcpyDeviceToDevice, ctx->stream));
		set_order(ctx, &S, samp);
		fsx_settings st{};
		st.maxiter = iters;
		st.rtol = 0.0f;
		st.use_zero_guess = 1;
		struct cand { int wg, mapping, lanes; };
		std::vector<cand> cands;
		for (int wg : {256, 512, 1024}) {
			if (has_sdia(B->A)) cands.push_back({wg, 2, 1});
			if (has_sliced(B->A)) cands.push_back({wg, 0, 1});
			for (int L : {2, 4, 8}) cands.push_back({wg, 1, L});
		}
		double best = 1e300;
		cand bc = {B->wg, B->mapping, B->lanes};
		hipEvent_t e0, e1;
		FSX_HIP(hipEventCreate(&e0));
		FSX_HIP(hipEventCreate(&e1));
		std::vector<int> it_out(m), st_out(m);
		for (const cand & c : cands) {
			S.wg = c.wg;
			S.mapping = c.mapping;
			S.lanes = c.lanes;
			// warm, then time
			for (int rep = 0; rep < 2; ++rep) {
				FSX_HIP(hipEventRecord(e0, ctx->stream));
				batched_cg_run(ctx, &S, work_for(B, false), dinv, b, x, &st, nullptr,
				               nullptr, nullptr);
				FSX_HIP(hipEventRecord(e1, ctx->stream));
				FSX_HIP(hipEventSynchronize(e1));
			}
			float ms = 0;
			FSX_HIP(hipEventElapsedTime(&ms, e0, e1));
			if (ms < best) {
				best = ms;
				bc = c;
			}
		}
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		B->wg = bc.wg;
		B->mapping = bc.mapping;
		B->lanes = bc.lanes;
		B->tuned = true;
		B->tuned_ms = best;
		cache[key] = std::make_tuple(bc.wg, bc.mapping, bc.lanes, best);
		if (best_ms) *best_ms = best;
	});
}

} // extern "C"
