// fsx_adjoint.hip -- discrete adjoint of backward-Euler diffusion trajectories.
//
// Forward: u^{s+1} = A^{-1} u^s, s = 0..steps-1, with A = I + gamma L_k the
// backward-Euler operator (time-integrators/operator_adapter.hh:29-35:
// y = x - gamma F(x); bdf.hh:292-296: gamma = dt, right-hand side the
// previous solution; bdf.hh:269-272: the previous solution is the initial
// guess).  Each step is one CG (+ Jacobi) solve through fsx_cg_solve, or
// (the _ir entry points) one mixed-precision refinement through
// fsx_ir_cg_solve on the symmetrically scaled FP32 copy of A.
//
// Loss L = |u^N - u_target|^2.  Adjoint: lambda^N = 2 (u^N - u_target);
// backwards, mu = A^{-T} lambda^{s+1}, dL/dA += -mu (u^{s+1})^T sampled on
// A's sparsity pattern, lambda^s = mu.  States are recomputed segment by
// segment from checkpoints kept every `ckpt` steps, so device memory holds
// ceil(steps/ckpt) + ckpt + 1 state vectors instead of steps + 1.
//
// The chain from dL/dA (values on the pattern) to the node coefficients of
// the 7-point operator shift I + scale L_k with arithmetic-mean faces
// k_f = (k_a + k_b)/2 (physics/volume_diffusion/coefficient.hh:157-174):
// dA/dk_f = scale (e_a - e_b)(e_a - e_b)^T for a face between two unknowns,
// scale e_a e_a^T for a face to a Dirichlet node; dL/dlog k_p = k_p sum over
// p's faces of dL/dk_f / 2.
#include "fsx_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace fsx {

// out[e] += (scale * lam[row]) * x[col] over the stored entries; 8 lanes per
// row, consecutive rows on consecutive lane groups (coalesced for short rows).
template<class I>
__global__ __launch_bounds__(256) void sampled_outer_kernel(
	int64_t nrows, const I * __restrict__ rp, const I * __restrict__ ci,
	const double * __restrict__ lam, const double * __restrict__ x, double scale,
	double * __restrict__ out) {
	const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const int64_t r = t >> 3;
	if (r >= nrows) return;
	const double a = scale * lam[r];
	for (I e = rp[r] + (I)(t & 7); e < rp[r + 1]; e += 8) out[e] += a * x[ci[e]];
}

// Lattice helpers for the 7-point operator on the interior of an (n+2)^3
// lattice: unknown r = i + n (j + n k) sits at node (i+1, j+1, k+1).
struct lattice {
	int n;
	int64_t L, L2;
	__device__ bool interior(int64_t a) const {
		const int64_t i = a % L, j = (a / L) % L, k = a / L2;
		return i >= 1 && i <= n && j >= 1 && j <= n && k >= 1 && k <= n;
	}
	__device__ int64_t row(int64_t a) const {
		const int64_t i = a % L, j = (a / L) % L, k = a / L2;
		return (i - 1) + (int64_t)n * ((j - 1) + (int64_t)n * (k - 1));
	}
	__device__ int64_t node(int64_t r) const {
		const int64_t i = r % n, j = (r / n) % n, k = r / ((int64_t)n * n);
		return (i + 1) + L * ((j + 1) + L * (k + 1));
	}
	// neighbour of a in direction d (-z,-y,-x,+x,+y,+z), or -1 off the lattice
	__device__ int64_t nb(int64_t a, int d) const {
		const int64_t i = a % L, j = (a / L) % L, k = a / L2;
		switch (d) {
			case 0: return k > 0 ? a - L2 : -1;
			case 1: return j > 0 ? a - L : -1;
			case 2: return i > 0 ? a - 1 : -1;
			case 3: return i < L - 1 ? a + 1 : -1;
			case 4: return j < L - 1 ? a + L : -1;
			default: return k < L - 1 ? a + L2 : -1;
		}
	}
};

// A's values = shift I + scale L_k, entry order -z,-y,-x,diag,+x,+y,+z (the
// order of fsx_gen_diffusion3d); diag = shift + scale * sum_d k_f(d) summed
// in direction order.  A row whose length does not match the stencil flags
// `bad` and is left unchanged.
template<class I>
__global__ __launch_bounds__(256) void diffusion3d_values_kernel(
	lattice g, const I * __restrict__ rp, const double * __restrict__ k,
	double shift, double scale, double * __restrict__ vals, unsigned * bad) {
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const int64_t N = (int64_t)g.n * g.n * g.n;
	if (r >= N) return;
	const int64_t a = g.node(r);
	double kf[6];
	bool in[6];
	double s = 0.0;
	int cnt = 1;
	for (int d = 0; d < 6; ++d) {
		const int64_t b = g.nb(a, d);
		kf[d] = d < 3 ? 0.5 * (k[a] + k[b]) : 0.5 * (k[b] + k[a]);
		s += kf[d];
		in[d] = g.interior(b);
		cnt += in[d];
	}
	I e = rp[r];
	if ((int64_t)(rp[r + 1] - e) != cnt) {
		atomicOr(bad, 1u);
		return;
	}
	for (int d = 0; d < 3; ++d)
		if (in[d]) vals[e++] = -scale * kf[d];
	vals[e++] = shift + scale * s;
	for (int d = 3; d < 6; ++d)
		if (in[d]) vals[e++] = -scale * kf[d];
}

// Position of column c in row r (rows of the stencil hold at most 7 entries).
template<class I>
__device__ __forceinline__ double entry(const I * rp, const I * ci, const double * G,
                                        int64_t r, int64_t c) {
	for (I e = rp[r]; e < rp[r + 1]; ++e)
		if ((int64_t)ci[e] == c) return G[e];
	return 0.0;
}

// grad[p] = k[p] * sum over p's faces (direction order) of (scale/2) * g_f,
// g_f = ((G_ll + G_hh) - G_lh) - G_hl for the face's lower node l and upper
// node h, absent terms (Dirichlet nodes) zero.  Faces between two Dirichlet
// nodes do not enter A and contribute nothing.
template<class I>
__global__ __launch_bounds__(256) void coeff_grad_kernel(
	lattice g, const I * __restrict__ rp, const I * __restrict__ ci,
	const double * __restrict__ G, const double * __restrict__ k, double scale,
	double * __restrict__ grad) {
	const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= g.L * g.L2) return;
	const double hs = 0.5 * scale;
	const bool pin = g.interior(p);
	double acc = 0.0;
	for (int d = 0; d < 6; ++d) {
		const int64_t q = g.nb(p, d);
		if (q < 0) continue;
		const bool qin = g.interior(q);
		if (!pin && !qin) continue;
		const int64_t lo = d < 3 ? q : p, hi = d < 3 ? p : q;
		const bool lin = d < 3 ? qin : pin, hin = d < 3 ? pin : qin;
		const int64_t rl = lin ? g.row(lo) : -1, rh = hin ? g.row(hi) : -1;
		const double gll = lin ? entry(rp, ci, G, rl, rl) : 0.0;
		const double ghh = hin ? entry(rp, ci, G, rh, rh) : 0.0;
		const double glh = lin && hin ? entry(rp, ci, G, rl, rh) : 0.0;
		const double ghl = lin && hin ? entry(rp, ci, G, rh, rl) : 0.0;
		acc += hs * (((gll + ghh) - glh) - ghl);
	}
	grad[p] = k[p] * acc;
}

template<class I, class V>
__global__ void gather_vals_kernel(int64_t nnz, const I * __restrict__ perm,
                                   const V * __restrict__ v, V * __restrict__ out) {
	const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e < nnz) out[e] = v[perm[e]];
}

} // namespace fsx

using namespace fsx;

namespace {

// Owned temporary vector (released on scope exit).
struct tmpvec {
	fsx_vec * v = nullptr;
	tmpvec(fsx_ctx * ctx, size_t n) {
		int rc = fsx_vec_create(ctx, n, FSX_F64, &v);
		if (rc) throw error(rc, get_error());
	}
	~tmpvec() {
		if (v) fsx_vec_release(v);
	}
	tmpvec(const tmpvec &) = delete;
	tmpvec(tmpvec && o) noexcept : v(o.v) { o.v = nullptr; }
};

void call(int rc) {
	if (rc) throw error(rc, get_error());
}

void check_diffusion_pattern(const fsx_csr * A, int n) {
	FSX_REQUIRE(n >= 1, "grid size must be >= 1");
	const int64_t N = (int64_t)n * n * n;
	FSX_REQUIRE(A->nrows == N && A->ncols == N &&
	                A->nnz == 7 * N - 6 * (int64_t)n * n,
	            "A does not have the 7-point pattern of an n^3 grid");
}

// Refined solves on A32 = S A S (S = D^{-1/2}) to the outer tolerance
// s->rtol: the restarted refinement (FP32 inner CG to inner.rtol, FP64
// outer residual and update) or, with rr, the FP32 CG with FP64 residual
// replacement (segments of inner.rtol reduction).
struct refine {
	const fsx_csr * A32;
	fsx_settings inner;
	bool rr = false;
};
constexpr int kRefineMaxOuter = 20;
constexpr int kRefineMaxSegments = 100;

// One backward-Euler step: x <- solution of A x = b from the guess x, by CG
// (+ Jacobi) or, with R, by mixed-precision refinement (inner iterations
// counted).
void be_step(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv,
             const fsx_vec * b, fsx_vec * x, const fsx_settings * s, int step,
             const char * what, int64_t & iters, const refine * R) {
	bool ok;
	if (R) {
		fsx_ir_info ii{};
		if (R->rr)
			call(fsx_ir_cg_rr_solve(ctx, A, R->A32, b, x, (double)s->rtol, kRefineMaxSegments,
			                        &R->inner, &ii));
		else
			call(fsx_ir_cg_solve(ctx, A, R->A32, 2, b, x, (double)s->rtol, kRefineMaxOuter,
			                     &R->inner, &ii));
		ok = ii.status == FSX_CONVERGED_RTOL;
		iters += ii.inner_iters_total;
	}
	else {
		fsx_info info{};
		call(fsx_cg_solve(ctx, A, dinv, b, x, s, &info));
		ok = info.status == FSX_CONVERGED_RTOL || info.status == FSX_CONVERGED_ATOL;
		iters += info.iters;
	}
	if (!ok)
		throw error(FSX_ERR_STATE, std::string(what) + " solve of step " +
		                               std::to_string(step) + " did not converge");
}

} // namespace

extern "C" {

int fsx_csr_transpose(fsx_ctx * ctx, const fsx_csr * A, fsx_csr ** out) {
	return guard([&] {
		FSX_REQUIRE(ctx && A && out, "null argument");
		const size_t isz = A->it == FSX_I32 ? 4 : 8, vsz = A->vt == FSX_F64 ? 8 : 4;
		// host structure of A^T: counting sort by column, rows ascending
		std::vector<int64_t> rp(A->nrows + 1), ci(A->nnz);
		{
			std::vector<char> hrp((A->nrows + 1) * isz), hci(std::max<int64_t>(1, A->nnz) * isz);
			FSX_HIP(hipMemcpyAsync(hrp.data(), A->rowptr(), hrp.size(), hipMemcpyDeviceToHost,
			                       ctx->stream));
			if (A->nnz)
				FSX_HIP(hipMemcpyAsync(hci.data(), A->colind(), A->nnz * isz,
				                       hipMemcpyDeviceToHost, ctx->stream));
			FSX_HIP(hipStreamSynchronize(ctx->stream));
			for (int64_t i = 0; i <= A->nrows; ++i)
				rp[i] = isz == 4 ? ((int32_t *)hrp.data())[i] : ((int64_t *)hrp.data())[i];
			for (int64_t e = 0; e < A->nnz; ++e)
				ci[e] = isz == 4 ? ((int32_t *)hci.data())[e] : ((int64_t *)hci.data())[e];
		}
		std::vector<int64_t> trp(A->ncols + 1, 0), tci(A->nnz), perm(A->nnz);
		for (int64_t e = 0; e < A->nnz; ++e) trp[ci[e] + 1]++;
		for (int64_t j = 0; j < A->ncols; ++j) trp[j + 1] += trp[j];
		std::vector<int64_t> fill(trp.begin(), trp.end() - 1);
		for (int64_t i = 0; i < A->nrows; ++i)
			for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
				const int64_t pos = fill[ci[e]]++;
				tci[pos] = i;
				perm[pos] = e;
			}
		// values of A^T gathered on the device
		dbuf dperm(std::max<int64_t>(1, A->nnz) * isz);
		dbuf tvals(std::max<int64_t>(1, A->nnz) * vsz);
		std::vector<char> hperm(std::max<int64_t>(1, A->nnz) * isz);
		for (int64_t e = 0; e < A->nnz; ++e) {
			if (isz == 4) ((int32_t *)hperm.data())[e] = (int32_t)perm[e];
			else ((int64_t *)hperm.data())[e] = perm[e];
		}
		if (A->nnz) {
			FSX_HIP(hipMemcpyAsync(dperm.p, hperm.data(), A->nnz * isz, hipMemcpyHostToDevice,
			                       ctx->stream));
			const unsigned g = (unsigned)((A->nnz + 255) / 256);
			if (isz == 4 && vsz == 8)
				hipLaunchKernelGGL((gather_vals_kernel<int32_t, double>), dim3(g), dim3(256), 0,
				                   ctx->stream, A->nnz, (const int32_t *)dperm.p,
				                   (const double *)A->vals->p, (double *)tvals.p);
			else if (isz == 4)
				hipLaunchKernelGGL((gather_vals_kernel<int32_t, float>), dim3(g), dim3(256), 0,
				                   ctx->stream, A->nnz, (const int32_t *)dperm.p,
				                   (const float *)A->vals->p, (float *)tvals.p);
			else if (vsz == 8)
				hipLaunchKernelGGL((gather_vals_kernel<int64_t, double>), dim3(g), dim3(256), 0,
				                   ctx->stream, A->nnz, (const int64_t *)dperm.p,
				                   (const double *)A->vals->p, (double *)tvals.p);
			else
				hipLaunchKernelGGL((gather_vals_kernel<int64_t, float>), dim3(g), dim3(256), 0,
				                   ctx->stream, A->nnz, (const int64_t *)dperm.p,
				                   (const float *)A->vals->p, (float *)tvals.p);
			FSX_HIP(hipGetLastError());
		}
		std::vector<char> hv(std::max<int64_t>(1, A->nnz) * vsz);
		if (A->nnz)
			FSX_HIP(hipMemcpyAsync(hv.data(), tvals.p, A->nnz * vsz, hipMemcpyDeviceToHost,
			                       ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		std::vector<char> htrp((A->ncols + 1) * isz), htci(std::max<int64_t>(1, A->nnz) * isz);
		for (int64_t j = 0; j <= A->ncols; ++j) {
			if (isz == 4) ((int32_t *)htrp.data())[j] = (int32_t)trp[j];
			else ((int64_t *)htrp.data())[j] = trp[j];
		}
		for (int64_t e = 0; e < A->nnz; ++e) {
			if (isz == 4) ((int32_t *)htci.data())[e] = (int32_t)tci[e];
			else ((int64_t *)htci.data())[e] = tci[e];
		}
		call(fsx_csr_create(ctx, A->ncols, A->nrows, A->nnz, A->it, htrp.data(), htci.data(),
		                    A->vt, hv.data(), out));
	});
}

int fsx_csr_sampled_outer(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * lam,
                          const fsx_vec * x, double scale, fsx_vec * out_vals) {
	return guard([&] {
		FSX_REQUIRE(ctx && A, "null argument");
		check_vec(lam, "lam");
		check_vec(x, "x");
		check_vec(out_vals, "out_vals");
		FSX_REQUIRE(lam->dt == FSX_F64 && x->dt == FSX_F64 && out_vals->dt == FSX_F64,
		            "sampled outer product is FP64");
		FSX_REQUIRE((int64_t)lam->n == A->nrows && (int64_t)x->n == A->ncols &&
		                (int64_t)out_vals->n == A->nnz,
		            "sampled outer: lam has nrows, x ncols, out_vals nnz entries");
		FSX_REQUIRE(out_vals->ptr() != lam->ptr() && out_vals->ptr() != x->ptr(),
		            "out_vals must not alias lam or x");
		if (A->nrows == 0) return;
		const unsigned g = (unsigned)((A->nrows * 8 + 255) / 256);
		if (A->it == FSX_I32)
			hipLaunchKernelGGL((sampled_outer_kernel<int32_t>), dim3(g), dim3(256), 0,
			                   ctx->stream, A->nrows, (const int32_t *)A->rowptr(),
			                   (const int32_t *)A->colind(), (const double *)lam->ptr(),
			                   (const double *)x->ptr(), scale, (double *)out_vals->ptr());
		else
			hipLaunchKernelGGL((sampled_outer_kernel<int64_t>), dim3(g), dim3(256), 0,
			                   ctx->stream, A->nrows, (const int64_t *)A->rowptr(),
			                   (const int64_t *)A->colind(), (const double *)lam->ptr(),
			                   (const double *)x->ptr(), scale, (double *)out_vals->ptr());
		FSX_HIP(hipGetLastError());
		ctx->launches++;
	});
}

int fsx_diffusion3d_values(fsx_ctx * ctx, int n, const fsx_vec * k, double shift,
                           double scale, fsx_csr * A) {
	return guard([&] {
		FSX_REQUIRE(ctx && A, "null argument");
		check_diffusion_pattern(A, n);
		check_vec(k, "k");
		FSX_REQUIRE(k->dt == FSX_F64 && (int64_t)k->n == (int64_t)(n + 2) * (n + 2) * (n + 2),
		            "k: FP64 on the (n+2)^3 lattice");
		FSX_REQUIRE(A->vt == FSX_F64, "A must hold FP64 values");
		lattice g{n, (int64_t)n + 2, ((int64_t)n + 2) * (n + 2)};
		dbuf bad(sizeof(unsigned));
		FSX_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned), ctx->stream));
		const int64_t N = (int64_t)n * n * n;
		const unsigned gr = (unsigned)((N + 255) / 256);
		if (A->it == FSX_I32)
			hipLaunchKernelGGL((diffusion3d_values_kernel<int32_t>), dim3(gr), dim3(256), 0,
			                   ctx->stream, g, (const int32_t *)A->rowptr(),
			                   (const double *)k->ptr(), shift, scale, (double *)A->vals->p,
			                   (unsigned *)bad.p);
		else
			hipLaunchKernelGGL((diffusion3d_values_kernel<int64_t>), dim3(gr), dim3(256), 0,
			                   ctx->stream, g, (const int64_t *)A->rowptr(),
			                   (const double *)k->ptr(), shift, scale, (double *)A->vals->p,
			                   (unsigned *)bad.p);
		FSX_HIP(hipGetLastError());
		ctx->launches++;
		refresh_sliced_values(ctx, A);
		unsigned hbad = 0;
		FSX_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(unsigned), hipMemcpyDeviceToHost,
		                       ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		FSX_REQUIRE(hbad == 0, "A's row lengths do not match the 7-point stencil");
	});
}

int fsx_diffusion3d_coeff_grad(fsx_ctx * ctx, int n, const fsx_csr * A, const fsx_vec * G,
                               const fsx_vec * k, double scale, fsx_vec * grad_logk) {
	return guard([&] {
		FSX_REQUIRE(ctx && A, "null argument");
		check_diffusion_pattern(A, n);
		check_vec(G, "G");
		check_vec(k, "k");
		check_vec(grad_logk, "grad_logk");
		const int64_t NL = (int64_t)(n + 2) * (n + 2) * (n + 2);
		FSX_REQUIRE(G->dt == FSX_F64 && (int64_t)G->n == A->nnz, "G: FP64, nnz entries");
		FSX_REQUIRE(k->dt == FSX_F64 && (int64_t)k->n == NL && grad_logk->dt == FSX_F64 &&
		                (int64_t)grad_logk->n == NL,
		            "k and grad_logk: FP64 on the (n+2)^3 lattice");
		FSX_REQUIRE(grad_logk->ptr() != G->ptr() && grad_logk->ptr() != k->ptr(),
		            "grad_logk must not alias G or k");
		lattice g{n, (int64_t)n + 2, ((int64_t)n + 2) * (n + 2)};
		const unsigned gr = (unsigned)((NL + 255) / 256);
		if (A->it == FSX_I32)
			hipLaunchKernelGGL((coeff_grad_kernel<int32_t>), dim3(gr), dim3(256), 0, ctx->stream,
			                   g, (const int32_t *)A->rowptr(), (const int32_t *)A->colind(),
			                   (const double *)G->ptr(), (const double *)k->ptr(), scale,
			                   (double *)grad_logk->ptr());
		else
			hipLaunchKernelGGL((coeff_grad_kernel<int64_t>), dim3(gr), dim3(256), 0, ctx->stream,
			                   g, (const int64_t *)A->rowptr(), (const int64_t *)A->colind(),
			                   (const double *)G->ptr(), (const double *)k->ptr(), scale,
			                   (double *)grad_logk->ptr());
		FSX_HIP(hipGetLastError());
		ctx->launches++;
	});
}

static void be_forward_impl(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv,
                            const fsx_vec * u0, int steps, const fsx_settings * s, fsx_vec * uT,
                            fsx_be_stats * st, const refine * R) {
	{
		FSX_REQUIRE(ctx && A && s, "null argument");
		check_vec(u0, "u0");
		check_vec(uT, "uT");
		FSX_REQUIRE(steps >= 0, "steps must be >= 0");
		FSX_REQUIRE(A->nrows == A->ncols && (int64_t)u0->n == A->nrows &&
		                (int64_t)uT->n == A->nrows && u0->dt == FSX_F64 && uT->dt == FSX_F64,
		            "u0, uT: FP64, one entry per row of the square A");
		fsx_be_stats S{};
		S.failed_step = -1;
		tmpvec a(ctx, A->nrows), b(ctx, A->nrows);
		fsx_vec * cur = a.v, * nxt = b.v;
		launch_copy(ctx, u0, cur);
		for (int k = 0; k < steps; ++k) {
			launch_copy(ctx, cur, nxt); // initial guess: the previous solution
			try {
				be_step(ctx, A, dinv, cur, nxt, s, k, "forward", S.forward_iters, R);
			}
			catch (...) {
				S.failed_step = k;
				if (st) *st = S;
				throw;
			}
			S.forward_solves++;
			std::swap(cur, nxt);
		}
		launch_copy(ctx, cur, uT);
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		if (st) *st = S;
	}
}

static void be_adjoint_impl(fsx_ctx * ctx, const fsx_csr * A, const fsx_csr * AT,
                            const fsx_vec * dinv, const fsx_vec * u0, const fsx_vec * u_target,
                            int steps, int ckpt, const fsx_settings * s, double * loss,
                            fsx_vec * G, fsx_vec * uT, fsx_vec * grad_u0, fsx_be_stats * st,
                            const refine * R) {
	{
		FSX_REQUIRE(ctx && A && s && loss, "null argument");
		check_vec(u0, "u0");
		check_vec(u_target, "u_target");
		check_vec(G, "G");
		FSX_REQUIRE(steps >= 1 && ckpt >= 1, "steps and ckpt must be >= 1");
		const int64_t N = A->nrows;
		FSX_REQUIRE(A->ncols == N && (int64_t)u0->n == N && (int64_t)u_target->n == N &&
		                u0->dt == FSX_F64 && u_target->dt == FSX_F64,
		            "u0, u_target: FP64, one entry per row of the square A");
		FSX_REQUIRE(G->dt == FSX_F64 && (int64_t)G->n == A->nnz, "G: FP64, nnz entries");
		if (AT)
			FSX_REQUIRE(AT->nrows == N && AT->ncols == N && AT->nnz == A->nnz,
			            "AT must be the transpose of A");
		if (uT) FSX_REQUIRE((int64_t)uT->n == N && uT->dt == FSX_F64, "uT: FP64, N entries");
		if (grad_u0)
			FSX_REQUIRE((int64_t)grad_u0->n == N && grad_u0->dt == FSX_F64,
			            "grad_u0: FP64, N entries");
		const fsx_csr * AA = AT ? AT : A;
		fsx_be_stats S{};
		S.failed_step = -1;
		auto fail_at = [&](int k) {
			S.failed_step = k;
			if (st) *st = S;
		};
		// forward pass keeping every ckpt-th state, and every state of the
		// last segment (its backward pass needs no recomputation)
		const int nck = (steps + ckpt - 1) / ckpt;
		const int last0 = (nck - 1) * ckpt;
		std::vector<tmpvec> cks, seg;
		cks.reserve(nck);
		seg.reserve(ckpt + 1);
		for (int c = 0; c < nck; ++c) cks.emplace_back(ctx, N);
		for (int i = 0; i <= ckpt; ++i) seg.emplace_back(ctx, N);
		S.checkpoint_bytes = (int64_t)(nck + ckpt + 1) * N * 8;
		tmpvec a(ctx, N), b(ctx, N);
		fsx_vec * cur = a.v, * nxt = b.v;
		launch_copy(ctx, u0, cur);
		for (int k = 0; k < steps; ++k) {
			if (k % ckpt == 0) launch_copy(ctx, cur, cks[k / ckpt].v);
			if (k == last0) launch_copy(ctx, cur, seg[0].v);
			launch_copy(ctx, cur, nxt);
			try {
				be_step(ctx, A, dinv, cur, nxt, s, k, "forward", S.forward_iters, R);
			}
			catch (...) {
				fail_at(k);
				throw;
			}
			S.forward_solves++;
			std::swap(cur, nxt);
			if (k >= last0) launch_copy(ctx, cur, seg[k - last0 + 1].v);
		}
		if (uT) launch_copy(ctx, cur, uT);
		// lambda^N = 2 (u^N - u_target), L = |u^N - u_target|^2
		tmpvec lam_a(ctx, N), lam_b(ctx, N);
		fsx_vec * lam = lam_a.v, * mu = lam_b.v;
		launch_axpy(ctx, -1.0, u_target, cur, lam); // lam = u^N - u_target
		*loss = reduce_host(ctx, RED_DOT, lam, lam);
		launch_axpy(ctx, 1.0, lam, lam, lam);        // lam = 2 (u^N - u_target)
		launch_set(ctx, 0.0, G);
		// backward over segments, recomputing the states of all but the last
		// (the solves are deterministic: recomputed states equal the forward's)
		for (int c = nck - 1; c >= 0; --c) {
			const int s0 = c * ckpt, s1 = std::min(steps, s0 + ckpt);
			if (c != nck - 1) launch_copy(ctx, cks[c].v, seg[0].v);
			for (int k = s0; k < s1 && c != nck - 1; ++k) {
				fsx_vec * x = seg[k - s0 + 1].v;
				launch_copy(ctx, seg[k - s0].v, x);
				try {
					be_step(ctx, A, dinv, seg[k - s0].v, x, s, k, "recompute",
					        S.recompute_iters, R);
				}
				catch (...) {
					fail_at(k);
					throw;
				}
				S.recompute_solves++;
			}
			for (int k = s1 - 1; k >= s0; --k) {
				// mu = A^{-T} lambda^{k+1}, guess lambda^{k+1}
				launch_copy(ctx, lam, mu);
				try {
					be_step(ctx, AA, dinv, lam, mu, s, k, "adjoint", S.adjoint_iters, R);
				}
				catch (...) {
					fail_at(k);
					throw;
				}
				S.adjoint_solves++;
				// dL/dA += -mu (u^{k+1})^T on the pattern
				call(fsx_csr_sampled_outer(ctx, A, mu, seg[k - s0 + 1].v, -1.0, G));
				std::swap(lam, mu);
			}
		}
		if (grad_u0) launch_copy(ctx, lam, grad_u0);
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		if (st) *st = S;
	}
}


int fsx_be_forward(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv, const fsx_vec * u0,
                   int steps, const fsx_settings * s, fsx_vec * uT, fsx_be_stats * st) {
	return guard([&] { be_forward_impl(ctx, A, dinv, u0, steps, s, uT, st, nullptr); });
}

int fsx_be_adjoint(fsx_ctx * ctx, const fsx_csr * A, const fsx_csr * AT, const fsx_vec * dinv,
                   const fsx_vec * u0, const fsx_vec * u_target, int steps, int ckpt,
                   const fsx_settings * s, double * loss, fsx_vec * G, fsx_vec * uT,
                   fsx_vec * grad_u0, fsx_be_stats * st) {
	return guard([&] {
		be_adjoint_impl(ctx, A, AT, dinv, u0, u_target, steps, ckpt, s, loss, G, uT, grad_u0, st,
		                nullptr);
	});
}

static refine refine_args(const fsx_csr * A, const fsx_csr * A32, const fsx_settings * inner,
                          bool rr = false) {
	FSX_REQUIRE(A && A32 && inner, "null argument");
	FSX_REQUIRE(A32->vt == FSX_F32 && A32->idx == A->idx,
	            "A32 must be fsx_csr_convert(A, FSX_F32) holding "
	            "fsx_csr_set_values_f32_sym_scaled values");
	if (rr)
		FSX_REQUIRE(inner->rtol > 0.0 && inner->rtol < 1.0 && inner->maxiter >= 1,
		            "the segment reduction inner->rtol must lie in (0, 1), maxiter >= 1");
	return refine{A32, *inner, rr};
}

int fsx_be_forward_ir(fsx_ctx * ctx, const fsx_csr * A, const fsx_csr * A32,
                      const fsx_vec * u0, int steps, const fsx_settings * s,
                      const fsx_settings * inner, fsx_vec * uT, fsx_be_stats * st) {
	return guard([&] {
		const refine R = refine_args(A, A32, inner);
		be_forward_impl(ctx, A, nullptr, u0, steps, s, uT, st, &R);
	});
}

int fsx_be_adjoint_ir(fsx_ctx * ctx, const fsx_csr * A, const fsx_csr * A32,
                      const fsx_vec * u0, const fsx_vec * u_target, int steps, int ckpt,
                      const fsx_settings * s, const fsx_settings * inner, double * loss,
                      fsx_vec * G, fsx_vec * uT, fsx_vec * grad_u0, fsx_be_stats * st) {
	return guard([&] {
		const refine R = refine_args(A, A32, inner);
		// A symmetric: the adjoint solves use A (and A32) themselves
		be_adjoint_impl(ctx, A, nullptr, nullptr, u0, u_target, steps, ckpt, s, loss, G, uT,
		                grad_u0, st, &R);
	});
}

int fsx_be_forward_rr(fsx_ctx * ctx, const fsx_csr * A, const fsx_csr * A32,
                      const fsx_vec * u0, int steps, const fsx_settings * s,
                      const fsx_settings * inner, fsx_vec * uT, fsx_be_stats * st) {
	return guard([&] {
		const refine R = refine_args(A, A32, inner, true);
		be_forward_impl(ctx, A, nullptr, u0, steps, s, uT, st, &R);
	});
}

int fsx_be_adjoint_rr(fsx_ctx * ctx, const fsx_csr * A, const fsx_csr * A32,
                      const fsx_vec * u0, const fsx_vec * u_target, int steps, int ckpt,
                      const fsx_settings * s, const fsx_settings * inner, double * loss,
                      fsx_vec * G, fsx_vec * uT, fsx_vec * grad_u0, fsx_be_stats * st) {
	return guard([&] {
		const refine R = refine_args(A, A32, inner, true);
		be_adjoint_impl(ctx, A, nullptr, nullptr, u0, u_target, steps, ckpt, s, loss, G, uT,
		                grad_u0, st, &R);
	});
}

} // extern "C"
