// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (built by
// tests/test_sanitize_host.py with g++ -fsanitize=address,undefined; no
// device use): the oracle's solvers, matrix readers and multigrid on small
// problems, and the library's host input path (fsx_io.cpp: Matrix Market
// header and entries, COO -> CSR, fsx_csr_from_mm) on the golden files and
// on malformed input.  fsx_csr_from_mm hands its CSR to fsx_csr_create; here
// that is a checking stand-in (monotone row pointers, columns in range)
// instead of the device upload.  argv[1]: tests/golden.  Exit 0 = pass.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "fsx.h"

extern "C" {
typedef struct {
	int maxiter;
	float rtol, atol;
	int use_zero_guess, max_krylov_dim, pre_side, restart;
} or_settings;
typedef struct {
	int status, iters, restarts;
	float res_norm_initial, res_norm_final, sol_norm_initial, sol_norm_final, rhs_norm;
} or_info;
typedef struct {
	int64_t nrows, ncols, nnz;
	int itype;
	const void * rowptr;
	const void * colind;
	const double * vals;
} or_csr;
typedef struct {
	int kind;
	const or_csr * M;
	float omega;
	int nrelax;
} or_prec;
typedef int (*or_diag_fn)(void *, const double *, int64_t, double);
typedef struct {
	int max_levels;
	int64_t min_coarse;
	int maxiter;
	float atol, beta;
	int pairwise_passes, nrelax, cycle, kappa;
	float ktol;
	int smoother;
	float jacobi_weight;
} or_amg_settings;
struct or_amg;

int or_cg(const or_csr *, const or_prec *, const double *, double *, const or_settings *,
          or_diag_fn, void *, or_info *, double *, int);
int or_fcg(const or_csr *, const or_prec *, const double *, double *, const or_settings *,
           or_diag_fn, void *, or_info *);
int or_gmres(const or_csr *, const or_prec *, const double *, double *, const or_settings *,
             or_diag_fn, void *, or_info *, double *, int);
int or_bicgstab(const or_csr *, const or_prec *, const double *, double *,
                const or_settings *, or_diag_fn, void *, or_info *, double *, int);
int or_cg_colours(const or_csr *, const double *, const double *, double *,
                  const or_settings *, int, or_info *);
void or_spmv(const or_csr *, const double *, double *);
void or_coo_tocsr(int64_t, int64_t, const int64_t *, const int64_t *, const double *,
                  int64_t *, int64_t *, double *);
void or_dinv_csr(const or_csr *, int64_t *, int64_t *, double *);
int64_t or_mm_read(const char *, int64_t *, int64_t *, int *, int64_t *, int64_t *, double *,
                   int64_t);
void or_jacobi_relax(const or_csr *, float, int, const double *, double *);
int or_hybrid_gs(const or_csr *, int, int, int, const double *, double *);
or_amg * or_amg_create(const or_csr *, const or_amg_settings *);
int or_amg_apply(or_amg *, const double *, double *, int *, double *);
void or_amg_destroy(or_amg *);
double or_l2norm(const double *, int64_t);
void or_set_random(double *, int64_t, unsigned);
}

// ---- stand-ins for the device side of fsx_io.cpp ---------------------------
namespace fsx {
static std::string g_msg;
void set_error(const std::string & m) { g_msg = m; }
}
static int g_created = 0;
extern "C" int fsx_csr_create(fsx_ctx *, int64_t nr, int64_t nc, int64_t nnz, fsx_itype it,
                              const void * rp, const void * ci, fsx_dtype vt, const void * v,
                              fsx_csr ** out) {
	if (vt != FSX_F64 || !out) return FSX_ERR_ARG;
	auto ok = [&](auto * r, auto * c) {
		if (r[0] != 0 || r[nr] != nnz) return false;
		for (int64_t i = 0; i < nr; ++i)
			if (r[i + 1] < r[i]) return false;
		for (int64_t e = 0; e < nnz; ++e)
			if (c[e] < 0 || c[e] >= nc) return false;
		double s = 0;
		for (int64_t e = 0; e < nnz; ++e) s += ((const double *)v)[e];
		return std::isfinite(s);
	};
	const bool good = it == FSX_I32 ? ok((const int32_t *)rp, (const int32_t *)ci)
	                                : ok((const int64_t *)rp, (const int64_t *)ci);
	if (!good) return FSX_ERR_ARG;
	++g_created;
	*out = nullptr;
	return FSX_OK;
}

static int fails = 0;
static void check(bool c, const char * what) {
	if (!c) {
		std::printf("FAIL %s\n", what);
		++fails;
	}
}

// 2D 5-point Laplacian on an n x n interior grid, int64 indices
static void lap2d(int n, std::vector<int64_t> & rp, std::vector<int64_t> & ci,
                  std::vector<double> & v) {
	rp.assign(1, 0);
	ci.clear();
	v.clear();
	for (int j = 0; j < n; ++j)
		for (int i = 0; i < n; ++i) {
			auto add = [&](int ii, int jj, double a) {
				if (ii >= 0 && ii < n && jj >= 0 && jj < n) {
					ci.push_back((int64_t)jj * n + ii);
					v.push_back(a);
				}
			};
			add(i, j - 1, -1);
			add(i - 1, j, -1);
			add(i, j, 4);
			add(i + 1, j, -1);
			add(i, j + 1, -1);
			rp.push_back((int64_t)ci.size());
		}
}

int main(int argc, char ** argv) {
	const std::string golden = argc > 1 ? argv[1] : "tests/golden";
	const int n = 20, N = n * n;
	std::vector<int64_t> rp, ci;
	std::vector<double> v;
	lap2d(n, rp, ci, v);
	or_csr A{N, N, (int64_t)ci.size(), 1, rp.data(), ci.data(), v.data()};
	std::vector<double> b(N), x(N), y(N);
	or_set_random(b.data(), N, 3);
	// Dinv as a CSR (the preconditioner of the reference's tests)
	std::vector<int64_t> drp(N + 1), dci(N);
	std::vector<double> dv(N);
	or_dinv_csr(&A, drp.data(), dci.data(), dv.data());
	or_csr D{N, N, N, 1, drp.data(), dci.data(), dv.data()};
	or_prec P{1, &D, 0.f, 0};
	or_prec J{2, &A, 2.f / 3.f, 2};
	or_settings s{500, 1e-10f, 0.f, 1, 30, 1, 1};
	or_info info;
	std::vector<double> hist(600);
	auto resid = [&](const std::vector<double> & xx) {
		or_spmv(&A, xx.data(), y.data());
		for (int i = 0; i < N; ++i) y[i] = b[i] - y[i];
		return or_l2norm(y.data(), N) / or_l2norm(b.data(), N);
	};
	for (const or_prec * pc : {(const or_prec *)nullptr, (const or_prec *)&P}) {
		x.assign(N, 0.0);
		or_cg(&A, pc, b.data(), x.data(), &s, nullptr, nullptr, &info, hist.data(), 600);
		check(info.status == 1 && resid(x) < 1.01e-10, "cg");
		x.assign(N, 0.0);
		or_fcg(&A, pc, b.data(), x.data(), &s, nullptr, nullptr, &info);
		check(info.status == 1 && resid(x) < 1.01e-10, "fcg");
		x.assign(N, 0.0);
		or_gmres(&A, pc, b.data(), x.data(), &s, nullptr, nullptr, &info, hist.data(), 600);
		check(info.status == 1 && resid(x) < 1e-9, "gmres");
		x.assign(N, 0.0);
		or_bicgstab(&A, pc, b.data(), x.data(), &s, nullptr, nullptr, &info, hist.data(), 600);
		check(info.status == 1 && resid(x) < 1e-9, "bicgstab");
	}
	x.assign(N, 0.0);
	or_cg(&A, &J, b.data(), x.data(), &s, nullptr, nullptr, &info, nullptr, 0);
	check(std::isfinite(resid(x)), "cg + weighted Jacobi");
	x.assign(N, 0.0);
	or_cg_colours(&A, dv.data(), b.data(), x.data(), &s, 4, &info);
	check(info.status == 1 && resid(x) < 1.01e-10, "cg over 4 colours");
	x.assign(N, 0.0);
	or_jacobi_relax(&A, 2.f / 3.f, 3, b.data(), x.data());
	check(or_hybrid_gs(&A, 3, 2, 2, b.data(), x.data()) == 0, "hybrid gs");
	or_amg_settings as{20, 5, 10, 1e-8f, 0.25f, 2, 1, 0, 0, 0.25f, 0, 2.f / 3.f};
	or_amg * M = or_amg_create(&A, &as);
	check(M != nullptr, "amg setup");
	if (M) {
		int cycles = 0;
		double res = 0;
		x.assign(N, 0.0);
		or_amg_apply(M, b.data(), x.data(), &cycles, &res);
		check(cycles > 0 && std::isfinite(res) && resid(x) < 0.5, "amg cycles");
		or_amg_destroy(M);
	}
	// COO -> CSR with duplicates and unsorted input
	{
		const int64_t I[] = {2, 0, 2, 1, 0, 2}, Jc[] = {0, 1, 2, 1, 1, 0};
		const double V[] = {1, 2, 3, 4, 5, 6};
		int64_t orp[4], oci[6], frp[4], fci[6];
		double ov[6], fv[6];
		or_coo_tocsr(3, 6, I, Jc, V, orp, oci, ov);
		check(fsx_coo_tocsr(3, 6, I, Jc, V, frp, fci, fv) == FSX_OK, "fsx_coo_tocsr");
		bool same = true;
		for (int i = 0; i < 4; ++i) same = same && orp[i] == frp[i];
		for (int e = 0; e < 6; ++e) same = same && oci[e] == fci[e] && ov[e] == fv[e];
		check(same, "tocsr agrees with the oracle");
	}
	// Matrix Market: the golden files through the oracle and the library
	for (const char * f : {"sym5.mtx", "gen4.mtx"}) {
		const std::string path = golden + "/" + f;
		int64_t nr, nc;
		int sym;
		const int64_t cap = or_mm_read(path.c_str(), &nr, &nc, &sym, nullptr, nullptr, nullptr, 0);
		check(cap > 0, "oracle mm header");
		std::vector<int64_t> I(cap), Jv(cap), fI(cap), fJ(cap);
		std::vector<double> V(cap), fV(cap);
		const int64_t got = or_mm_read(path.c_str(), &nr, &nc, &sym, I.data(), Jv.data(),
		                               V.data(), cap);
		int64_t fnr, fnc, lines, fcap, cnt = 0;
		int fsym;
		check(fsx_mm_read_header(path.c_str(), &fnr, &fnc, &lines, &fsym, &fcap) == FSX_OK,
		      "fsx mm header");
		check(fcap <= cap && fsx_mm_read_coo(path.c_str(), cap, fI.data(), fJ.data(),
		                                     fV.data(), &cnt) == FSX_OK,
		      "fsx mm entries");
		bool same = cnt == got && fnr == nr && fnc == nc;
		for (int64_t e = 0; same && e < cnt; ++e)
			same = I[e] == fI[e] && Jv[e] == fJ[e] && V[e] == fV[e];
		check(same, "mm reader agrees with the oracle");
		fsx_csr * out = nullptr;
		const int before = g_created;
		check(fsx_csr_from_mm(reinterpret_cast<fsx_ctx *>(&fails), path.c_str(), FSX_I32,
		                      &out) == FSX_OK &&
		          g_created == before + 1,
		      "fsx_csr_from_mm");
	}
	// malformed input: a missing file, fewer entry lines than announced,
	// too small a buffer, an out-of-range column
	{
		const char * tmp = std::getenv("TMPDIR") ? std::getenv("TMPDIR") : "/tmp";
		const std::string bad = std::string(tmp) + "/fsx_sanitize_bad.mtx";
		int64_t nr, nc, lines, cap, cnt;
		int sym;
		check(fsx_mm_read_header("/nonexistent/file.mtx", &nr, &nc, &lines, &sym, &cap) != FSX_OK,
		      "missing file rejected");
		FILE * f = std::fopen(bad.c_str(), "w");
		std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n3 3 4\n1 1 1.0\n2 2 2.0\n");
		std::fclose(f);
		std::vector<int64_t> I(8), Jv(8);
		std::vector<double> V(8);
		const int h = fsx_mm_read_header(bad.c_str(), &nr, &nc, &lines, &sym, &cap);
		check(h != FSX_OK || fsx_mm_read_coo(bad.c_str(), 8, I.data(), Jv.data(), V.data(),
		                                     &cnt) != FSX_OK,
		      "short file rejected");
		f = std::fopen(bad.c_str(), "w");
		std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n2 2 2.0\n1 2 3.0\n");
		std::fclose(f);
		check(fsx_mm_read_coo(bad.c_str(), 2, I.data(), Jv.data(), V.data(), &cnt) != FSX_OK,
		      "small buffer rejected");
		f = std::fopen(bad.c_str(), "w");
		std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n2 3 2.0\n");
		std::fclose(f);
		fsx_csr * out = nullptr;
		check(fsx_csr_from_mm(reinterpret_cast<fsx_ctx *>(&fails), bad.c_str(), FSX_I32, &out) !=
		          FSX_OK,
		      "out-of-range column rejected");
		std::remove(bad.c_str());
	}
	std::printf(fails ? "FAIL\n" : "PASS\n");
	return fails ? 1 : 0;
}
