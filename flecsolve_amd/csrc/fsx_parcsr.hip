// fsx_parcsr.hip -- the parallel CSR matrix (topo::csr colouring and
// mat::parcsr apply) with all of its colours on one device.
//
// The reference distributes a matrix over P colours (topo/csr.hh:482-618,
// matrices/parcsr.hh:100-172).  Rows and columns are cut into FleCSI
// equal_map blocks (block c starts at c q + min(c, m), q = n / P,
// m = n % P); colour c owns row block c and column block c.  init_mats
// (csr.hh:553-618) walks each owned row in CSR order and sends every entry
// either to the colour's diag block (its column lies in the colour's column
// block; stored with the local column id) or to its offd block (stored with
// ncols_local + the column's rank in the colour's ghost list, the sorted
// unique ghost columns of color(), csr.hh:509-523).  parcsr_ops::spmv
// (parcsr.hh:60-68) then computes
//     tmp = offd x_ghost     (spmv_remote, after the ghost copy of x)
//     y   = diag x_local     (spmv_local)
//     y   = y + tmp          (y.add(y, tmp))
// with each block's product summed from 0 in the row's CSR order
// (compressed_ops::spmv, matrices/seq.hh:177-193).  When a row has entries
// in both blocks the result differs in rounding from the one-colour product,
// so a P-colour run is its own reference result; this module reproduces it
// bit for bit.
//
// Device layout (every colour handled by the same launches):
//   D        the diag entries of all rows as one CSR (nrows x ncols, columns
//            global = local id + first column of the colour), multiplied by
//            the library's SpMV (any row mapping that sums in CSR order);
//   colmap   the ghost lists of all colours concatenated (global column ids);
//            colour c's list starts at gbase[c];
//   ghost    the ghost values, gathered from x through colmap: the copy plan
//            of csr.hh:276-300 (a halo exchange between devices, a gather on
//            one);
//   brow, bptr, bcol, bval
//            the rows that have offd entries, ascending, with their offd
//            entries in CSR order, columns indexing `ghost`.
// A row without offd entries keeps y = diag x unchanged: the reference adds
// tmp = 0 there, and y + 0 == y because a sum that starts from +0 is never
// -0.
#include "fsx_internal.hpp"

#include <algorithm>
#include <string>
#include <vector>

namespace fsx {

namespace {

// ghost[k] = x[colmap[k]]
template<class I>
__global__ __launch_bounds__(256) void ghost_gather_kernel(const I * __restrict__ colmap,
                                                           int64_t ng,
                                                           const double * __restrict__ x,
                                                           double * __restrict__ ghost) {
	const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (k >= ng) return;
	ghost[k] = x[colmap[k]];
}

// For each row with offd entries: tmp = sum in CSR order from 0, y = y + tmp.
template<class I>
__global__ __launch_bounds__(256) void offd_add_kernel(const I * __restrict__ brow,
                                                       const I * __restrict__ bptr,
                                                       const I * __restrict__ bcol,
                                                       const double * __restrict__ bval,
                                                       int64_t nb,
                                                       const double * __restrict__ ghost,
                                                       double * __restrict__ y) {
	const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (t >= nb) return;
	double tmp = 0.;
	const int64_t e1 = bptr[t + 1];
	for (int64_t e = bptr[t]; e < e1; ++e) tmp += bval[e] * ghost[bcol[e]];
	const int64_t r = brow[t];
	y[r] = y[r] + tmp;
}

// equal_map(n, P) block bounds
struct block_map {
	int64_t n, P, q, m;
	block_map(int64_t n_, int64_t P_) : n(n_), P(P_), q(n_ / P_), m(n_ % P_) {}
	int64_t lo(int64_t c) const { return c * q + std::min(c, m); }
};

void rc(int code) {
	if (code) throw error(code, get_error());
}

} // namespace

} // namespace fsx

using namespace fsx;

struct fsx_parcsr {
	fsx_ctx * ctx = nullptr;
	int64_t nrows = 0, ncols = 0;
	int P = 1;
	fsx_itype it = FSX_I32;
	fsx_csr * D = nullptr;
	int64_t ng = 0, nb = 0, nnz_offd = 0;
	std::unique_ptr<dbuf> colmap, ghost, brow, bptr, bcol, bval;
	// per colour: ghost list start (P + 1 entries), diag and offd entry
	// counts
	std::vector<int64_t> gbase, nnz_diag_c, nnz_offd_c;
	~fsx_parcsr() {
		if (D) fsx_csr_destroy(D);
	}
};

namespace {

template<class I>
std::unique_ptr<dbuf> upload_as(fsx_ctx * ctx, const std::vector<int64_t> & h) {
	std::vector<I> t(h.begin(), h.end());
	auto d = std::make_unique<dbuf>(std::max<size_t>(1, t.size()) * sizeof(I));
	if (!t.empty())
		FSX_HIP(hipMemcpyAsync(d->p, t.data(), t.size() * sizeof(I), hipMemcpyHostToDevice,
		                       ctx->stream));
	return d;
}

template<class I>
std::vector<int64_t> download_as(fsx_ctx * ctx, const dbuf & d, int64_t n) {
	std::vector<I> t((size_t)n);
	if (n)
		FSX_HIP(hipMemcpyAsync(t.data(), d.p, (size_t)n * sizeof(I), hipMemcpyDeviceToHost,
		                       ctx->stream));
	FSX_HIP(hipStreamSynchronize(ctx->stream));
	return std::vector<int64_t>(t.begin(), t.end());
}

std::vector<int64_t> download_idx(fsx_ctx * ctx, fsx_itype it, const dbuf & d, int64_t n) {
	return it == FSX_I32 ? download_as<int32_t>(ctx, d, n) : download_as<int64_t>(ctx, d, n);
}

std::unique_ptr<dbuf> upload_idx(fsx_ctx * ctx, fsx_itype it, const std::vector<int64_t> & h) {
	return it == FSX_I32 ? upload_as<int32_t>(ctx, h) : upload_as<int64_t>(ctx, h);
}

} // namespace

extern "C" {

int fsx_parcsr_create(fsx_ctx * ctx, const fsx_csr * A, int ncolours, fsx_parcsr ** out) {
	return guard([&] {
		FSX_REQUIRE(ctx, "null context");
		FSX_REQUIRE(A, "null matrix");
		FSX_REQUIRE(out, "null output");
		FSX_REQUIRE(A->vt == FSX_F64, "parcsr: the matrix must hold FP64 values");
		FSX_REQUIRE(ncolours >= 1 && ncolours <= A->nrows && ncolours <= A->ncols,
		            "parcsr: ncolours must be in [1, min(nrows, ncols)]");
		const int64_t n = A->nrows, nc = A->ncols, nnz = A->nnz;
		const fsx_itype it = A->it;
		const size_t isz = it == FSX_I32 ? 4 : 8;

		// the matrix as the colours' input rows (parcsr read_mat hands each
		// colour its row block of the global matrix)
		std::vector<char> hrp((size_t)(n + 1) * isz), hci((size_t)std::max<int64_t>(1, nnz) * isz);
		std::vector<double> hv((size_t)nnz);
		FSX_HIP(hipMemcpyAsync(hrp.data(), A->rowptr(), hrp.size(), hipMemcpyDeviceToHost,
		                       ctx->stream));
		if (nnz) {
			FSX_HIP(hipMemcpyAsync(hci.data(), A->colind(), (size_t)nnz * isz,
			                       hipMemcpyDeviceToHost, ctx->stream));
			FSX_HIP(hipMemcpyAsync(hv.data(), A->vals->p, (size_t)nnz * sizeof(double),
			                       hipMemcpyDeviceToHost, ctx->stream));
		}
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		auto rp = [&](int64_t i) -> int64_t {
			return it == FSX_I32 ? ((const int32_t *)hrp.data())[i]
			                     : ((const int64_t *)hrp.data())[i];
		};
		auto ci = [&](int64_t e) -> int64_t {
			return it == FSX_I32 ? ((const int32_t *)hci.data())[e]
			                     : ((const int64_t *)hci.data())[e];
		};

		auto M = std::make_unique<fsx_parcsr>();
		M->ctx = ctx;
		M->nrows = n;
		M->ncols = nc;
		M->P = ncolours;
		M->it = it;
		const block_map rm(n, ncolours), cm(nc, ncolours);

		std::vector<int64_t> drp((size_t)n + 1, 0), dci, colmap, brow, bptr{0}, bcol;
		std::vector<double> dv, bval;
		dci.reserve((size_t)nnz);
		dv.reserve((size_t)nnz);
		M->gbase.assign((size_t)ncolours + 1, 0);
		M->nnz_diag_c.assign((size_t)ncolours, 0);
		M->nnz_offd_c.assign((size_t)ncolours, 0);
		std::vector<int64_t> ghosts;
		for (int c = 0; c < ncolours; ++c) {
			const int64_t r0 = rm.lo(c), r1 = rm.lo(c + 1);
			const int64_t c0 = cm.lo(c), c1 = cm.lo(c + 1);
			// ghost list: sorted unique offd columns (color, csr.hh:509-523)
			ghosts.clear();
			for (int64_t r = r0; r < r1; ++r)
				for (int64_t e = rp(r); e < rp(r + 1); ++e) {
					const int64_t j = ci(e);
					if (j < c0 || j >= c1) ghosts.push_back(j);
				}
			std::sort(ghosts.begin(), ghosts.end());
			ghosts.erase(std::unique(ghosts.begin(), ghosts.end()), ghosts.end());
			const int64_t gb = (int64_t)colmap.size();
			colmap.insert(colmap.end(), ghosts.begin(), ghosts.end());
			M->gbase[c + 1] = (int64_t)colmap.size();
			// split each row in CSR order (init_mats, csr.hh:579-616)
			for (int64_t r = r0; r < r1; ++r) {
				bool has_offd = false;
				for (int64_t e = rp(r); e < rp(r + 1); ++e) {
					const int64_t j = ci(e);
					if (j >= c0 && j < c1) {
						dci.push_back(j);
						dv.push_back(hv[(size_t)e]);
						++M->nnz_diag_c[c];
					}
					else {
						const int64_t k =
							std::lower_bound(ghosts.begin(), ghosts.end(), j) - ghosts.begin();
						bcol.push_back(gb + k);
						bval.push_back(hv[(size_t)e]);
						++M->nnz_offd_c[c];
						has_offd = true;
					}
				}
				drp[(size_t)r + 1] = (int64_t)dci.size();
				if (has_offd) {
					brow.push_back(r);
					bptr.push_back((int64_t)bcol.size());
				}
			}
		}
		M->ng = (int64_t)colmap.size();
		M->nb = (int64_t)brow.size();
		M->nnz_offd = (int64_t)bcol.size();

		// diag block of every colour: one CSR with global column ids
		if (it == FSX_I32) {
			std::vector<int32_t> rp32(drp.begin(), drp.end()), ci32(dci.begin(), dci.end());
			rc(fsx_csr_create(ctx, n, nc, (int64_t)dci.size(), FSX_I32, rp32.data(),
			                  ci32.data(), FSX_F64, dv.data(), &M->D));
		}
		else
			rc(fsx_csr_create(ctx, n, nc, (int64_t)dci.size(), FSX_I64, drp.data(), dci.data(),
			                  FSX_F64, dv.data(), &M->D));
		// the diag product must sum each row serially in CSR order: the
		// vector-per-row mapping (kind 1) splits rows over lanes
		int kind = 0, lanes = 0;
		rc(fsx_csr_plan(M->D, &kind, &lanes));
		if (kind == 1) rc(fsx_csr_set_plan(M->D, 0, 0));

		M->colmap = upload_idx(ctx, it, colmap);
		M->brow = upload_idx(ctx, it, brow);
		M->bptr = upload_idx(ctx, it, bptr);
		M->bcol = upload_idx(ctx, it, bcol);
		M->bval = std::make_unique<dbuf>(std::max<size_t>(1, bval.size()) * sizeof(double));
		if (!bval.empty())
			FSX_HIP(hipMemcpyAsync(M->bval->p, bval.data(), bval.size() * sizeof(double),
			                       hipMemcpyHostToDevice, ctx->stream));
		M->ghost = std::make_unique<dbuf>(std::max<size_t>(1, colmap.size()) * sizeof(double));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		ctx_retain(ctx);
		*out = M.release();
	});
}

int fsx_parcsr_destroy(fsx_parcsr * M) {
	return guard([&] {
		if (!M) return;
		fsx_ctx * c = M->ctx;
		(void)hipStreamSynchronize(c->stream);
		delete M;
		ctx_release(c);
	});
}

int fsx_parcsr_colour_info(const fsx_parcsr * M, int colour, int64_t * sizes) {
	return guard([&] {
		FSX_REQUIRE(M, "null matrix");
		FSX_REQUIRE(sizes, "null output");
		FSX_REQUIRE(colour >= 0 && colour < M->P, "colour out of range");
		const block_map rm(M->nrows, M->P), cm(M->ncols, M->P);
		sizes[0] = rm.lo(colour + 1) - rm.lo(colour);
		sizes[1] = cm.lo(colour + 1) - cm.lo(colour);
		sizes[2] = M->gbase[colour + 1] - M->gbase[colour];
		sizes[3] = M->nnz_diag_c[colour];
		sizes[4] = M->nnz_offd_c[colour];
	});
}

int fsx_parcsr_download(fsx_ctx * ctx, const fsx_parcsr * M, int colour, int64_t * diag_rowptr,
                        int64_t * diag_colind, double * diag_vals, int64_t * offd_rowptr,
                        int64_t * offd_colind, double * offd_vals, int64_t * colmap) {
	return guard([&] {
		FSX_REQUIRE(ctx, "null context");
		FSX_REQUIRE(M, "null matrix");
		FSX_REQUIRE(colour >= 0 && colour < M->P, "colour out of range");
		const block_map rm(M->nrows, M->P), cm(M->ncols, M->P);
		const int64_t r0 = rm.lo(colour), r1 = rm.lo(colour + 1), c0 = cm.lo(colour);
		const int64_t nloc = cm.lo(colour + 1) - c0, gb = M->gbase[colour];
		// diag block: rows r0..r1 of D, columns made local
		const int64_t dnnz = M->D->nnz;
		const size_t isz = M->it == FSX_I32 ? 4 : 8;
		std::vector<char> hrp((size_t)(M->nrows + 1) * isz),
			hci((size_t)std::max<int64_t>(1, dnnz) * isz);
		std::vector<double> hv((size_t)dnnz);
		rc(fsx_csr_download(ctx, M->D, hrp.data(), hci.data(), hv.data()));
		auto rp = [&](int64_t i) -> int64_t {
			return isz == 4 ? ((const int32_t *)hrp.data())[i] : ((const int64_t *)hrp.data())[i];
		};
		auto ci = [&](int64_t e) -> int64_t {
			return isz == 4 ? ((const int32_t *)hci.data())[e] : ((const int64_t *)hci.data())[e];
		};
		const int64_t base = rp(r0);
		for (int64_t r = r0; r <= r1; ++r)
			if (diag_rowptr) diag_rowptr[r - r0] = rp(r) - base;
		for (int64_t e = base; e < rp(r1); ++e) {
			if (diag_colind) diag_colind[e - base] = ci(e) - c0;
			if (diag_vals) diag_vals[e - base] = hv[(size_t)e];
		}
		// offd block: the colour's rows among brow, columns renumbered after
		// the local ones
		const auto brow = download_idx(ctx, M->it, *M->brow, M->nb);
		const auto bptr = download_idx(ctx, M->it, *M->bptr, M->nb + 1);
		const auto bcol = download_idx(ctx, M->it, *M->bcol, M->nnz_offd);
		std::vector<double> bval((size_t)M->nnz_offd);
		if (M->nnz_offd)
			FSX_HIP(hipMemcpyAsync(bval.data(), M->bval->p, bval.size() * sizeof(double),
			                       hipMemcpyDeviceToHost, ctx->stream));
		FSX_HIP(hipStreamSynchronize(ctx->stream));
		int64_t t = std::lower_bound(brow.begin(), brow.end(), r0) - brow.begin();
		int64_t w = 0;
		if (offd_rowptr) offd_rowptr[0] = 0;
		for (int64_t r = r0; r < r1; ++r) {
			if (t < M->nb && brow[(size_t)t] == r) {
				for (int64_t e = bptr[(size_t)t]; e < bptr[(size_t)t + 1]; ++e, ++w) {
					if (offd_colind) offd_colind[w] = nloc + (bcol[(size_t)e] - gb);
					if (offd_vals) offd_vals[w] = bval[(size_t)e];
				}
				++t;
			}
			if (offd_rowptr) offd_rowptr[r - r0 + 1] = w;
		}
		if (colmap) {
			const int64_t ngc = M->gbase[colour + 1] - gb;
			const auto cmh = download_idx(ctx, M->it, *M->colmap, M->ng);
			for (int64_t k = 0; k < ngc; ++k) colmap[k] = cmh[(size_t)(gb + k)];
		}
	});
}

int fsx_parcsr_spmv(fsx_ctx * ctx, const fsx_parcsr * M, const fsx_vec * x, fsx_vec * y) {
	return guard([&] {
		FSX_REQUIRE(ctx, "null context");
		FSX_REQUIRE(M, "null matrix");
		// the ghost buffer is M's own, ordered on the stream of M's context
		FSX_REQUIRE(ctx == M->ctx, "parcsr spmv: use the context the matrix was created on");
		check_vec(x, "x");
		check_vec(y, "y");
		FSX_REQUIRE(x->dt == FSX_F64 && y->dt == FSX_F64, "parcsr spmv: FP64 vectors");
		FSX_REQUIRE((int64_t)x->n == M->ncols, "x length != ncols");
		FSX_REQUIRE((int64_t)y->n == M->nrows, "y length != nrows");
		FSX_REQUIRE(x->ptr() != y->ptr(), "spmv: x and y must be distinct vectors");
		const double * xp = (const double *)x->ptr();
		double * yp = (double *)y->ptr();
		double * gp = (double *)M->ghost->p;
		// ghost copy first (spmv_remote reads x's ghosts), then diag, then
		// the offd rows' sums added in
		if (M->ng > 0) {
			const unsigned g = (unsigned)((M->ng + 255) / 256);
			if (M->it == FSX_I32)
				hipLaunchKernelGGL((ghost_gather_kernel<int32_t>), dim3(g), dim3(256), 0,
				                   ctx->stream, (const int32_t *)M->colmap->p, M->ng, xp, gp);
			else
				hipLaunchKernelGGL((ghost_gather_kernel<int64_t>), dim3(g), dim3(256), 0,
				                   ctx->stream, (const int64_t *)M->colmap->p, M->ng, xp, gp);
			FSX_HIP(hipGetLastError());
		}
		rc(fsx_csr_spmv(ctx, M->D, x, y));
		if (M->nb > 0) {
			const unsigned g = (unsigned)((M->nb + 255) / 256);
			if (M->it == FSX_I32)
				hipLaunchKernelGGL((offd_add_kernel<int32_t>), dim3(g), dim3(256), 0, ctx->stream,
				                   (const int32_t *)M->brow->p, (const int32_t *)M->bptr->p,
				                   (const int32_t *)M->bcol->p, (const double *)M->bval->p,
				                   M->nb, (const double *)gp, yp);
			else
				hipLaunchKernelGGL((offd_add_kernel<int64_t>), dim3(g), dim3(256), 0, ctx->stream,
				                   (const int64_t *)M->brow->p, (const int64_t *)M->bptr->p,
				                   (const int64_t *)M->bcol->p, (const double *)M->bval->p,
				                   M->nb, (const double *)gp, yp);
			FSX_HIP(hipGetLastError());
		}
	});
}

int fsx_parcsr_residual(fsx_ctx * ctx, const fsx_parcsr * M, const fsx_vec * b,
                        const fsx_vec * x, fsx_vec * r) {
	return guard([&] {
		FSX_REQUIRE(ctx, "null context");
		check_vec(b, "b");
		check_vec(r, "r");
		FSX_REQUIRE(b->ptr() != r->ptr(), "residual: b and r must be distinct vectors");
		rc(fsx_parcsr_spmv(ctx, M, x, r));
		rc(fsx_subtract(ctx, b, r, r));
	});
}

int fsx_time_parcsr_spmv(fsx_ctx * ctx, const fsx_parcsr * M, const fsx_vec * x,
                         fsx_vec * y, int reps, double * avg_ms) {
	return guard([&] {
		FSX_REQUIRE(ctx && avg_ms, "null argument");
		FSX_REQUIRE(reps > 0, "reps must be > 0");
		rc(fsx_parcsr_spmv(ctx, M, x, y)); // argument checks before timing
		hipEvent_t e0, e1;
		FSX_HIP(hipEventCreate(&e0));
		FSX_HIP(hipEventCreate(&e1));
		FSX_HIP(hipEventRecord(e0, ctx->stream));
		for (int i = 0; i < reps; ++i) rc(fsx_parcsr_spmv(ctx, M, x, y));
		FSX_HIP(hipEventRecord(e1, ctx->stream));
		FSX_HIP(hipEventSynchronize(e1));
		float ms = 0;
		FSX_HIP(hipEventElapsedTime(&ms, e0, e1));
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		*avg_ms = (double)ms / reps;
	});
}

int fsx_parcsr_diag(const fsx_parcsr * M, fsx_csr ** D) {
	return guard([&] {
		FSX_REQUIRE(M, "null matrix");
		FSX_REQUIRE(D, "null output");
		*D = M->D;
	});
}

} // extern "C"
