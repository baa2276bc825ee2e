// fsx_io.cpp -- host input path: Matrix Market reader and COO -> CSR.
//
// Reader: matrix_market::read_header / read / parse_entry
// (flecsolve/matrices/io/matrix_market.hh:39-103, 157-179):
//   - header lines are read until the first line not starting with '%';
//     a line containing "symmetric" anywhere up to and including that size
//     line marks the matrix symmetric;
//   - the size line and every entry line are split on single spaces; sizes
//     go through atoi, row and column through stoi minus one, the value
//     through stof (rounded to float, then stored as double);
//   - exactly `nnz` entry lines follow; for a symmetric file the mirrored
//     entry (col, row, val) is appended right after each off-diagonal one.
// COO -> CSR: coo::tocsr (flecsolve/matrices/seq.hh:324-352): entries in
// stable row order (input order kept inside a row), duplicates kept, columns
// not sorted.  A counting sort by row (count, prefix, scatter in input order)
// gives exactly the stable sort's order in O(nnz + nrows).
#include "fsx.h"

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace fsx {
void set_error(const std::string & msg); // thread-local message of fsx_last_error()
}

namespace {

struct mm_header {
	int64_t nrows = 0, ncols = 0, nnz = 0;
	bool symmetric = false;
};

bool read_header(std::ifstream & fh, mm_header & h) {
	std::string line;
	int64_t sizes[3] = {0, 0, 0};
	while (std::getline(fh, line)) {
		if (line.find("symmetric") != std::string::npos) h.symmetric = true;
		if (line.empty() || line[0] != '%') {
			std::istringstream iss(line);
			for (int i = 0; i < 3; ++i) {
				std::string tok;
				std::getline(iss, tok, ' ');
				sizes[i] = std::atoi(tok.c_str());
			}
			h.nrows = sizes[0];
			h.ncols = sizes[1];
			h.nnz = sizes[2];
			return true;
		}
	}
	return false;
}

// One entry line: three single-space separated tokens.
bool parse_entry(const std::string & line, int64_t & row, int64_t & col, double & val) {
	std::istringstream iss(line);
	std::string tok[3];
	for (int i = 0; i < 3; ++i) std::getline(iss, tok[i], ' ');
	try {
		row = (int64_t)std::stoi(tok[0]) - 1;
		col = (int64_t)std::stoi(tok[1]) - 1;
		val = (double)std::stof(tok[2]);
	}
	catch (const std::exception &) {
		return false;
	}
	return true;
}

int fail(int code, const std::string & msg) {
	fsx::set_error(msg);
	return code;
}

// No C++ exception crosses the C ABI.
template<class F>
int io_guard(F && body) {
	try {
		return body();
	}
	catch (const std::bad_alloc &) {
		return fail(FSX_ERR_ALLOC, "host allocation failed");
	}
	catch (const std::exception & e) {
		return fail(FSX_ERR_STATE, e.what());
	}
	catch (...) {
		return fail(FSX_ERR_STATE, "unknown error");
	}
}

} // namespace

extern "C" {

int fsx_mm_read_header(const char * path, int64_t * nrows, int64_t * ncols,
                       int64_t * lines, int * symmetric, int64_t * entries) {
	return io_guard([&]() -> int {
		if (!path) return fail(FSX_ERR_ARG, "null path");
		std::ifstream fh(path);
		if (!fh) return fail(FSX_ERR_ARG, std::string("cannot open ") + path);
		mm_header h;
		if (!read_header(fh, h)) return fail(FSX_ERR_ARG, "no size line");
		if (h.nrows < 0 || h.ncols < 0 || h.nnz < 0) return fail(FSX_ERR_ARG, "negative size");
		if (nrows) *nrows = h.nrows;
		if (ncols) *ncols = h.ncols;
		if (lines) *lines = h.nnz;
		if (symmetric) *symmetric = h.symmetric ? 1 : 0;
		if (entries) *entries = h.symmetric ? 2 * h.nnz : h.nnz; // upper bound
		return FSX_OK;
	});
}

int fsx_mm_read_coo(const char * path, int64_t cap, int64_t * I, int64_t * J, double * V,
                    int64_t * count) {
	return io_guard([&]() -> int {
		if (!path || !I || !J || !V || !count) return fail(FSX_ERR_ARG, "null argument");
		std::ifstream fh(path);
		if (!fh) return fail(FSX_ERR_ARG, std::string("cannot open ") + path);
		mm_header h;
		if (!read_header(fh, h)) return fail(FSX_ERR_ARG, "no size line");
		int64_t cnt = 0;
		for (int64_t i = 0; i < h.nnz; ++i) {
			std::string line;
			if (!std::getline(fh, line))
				return fail(FSX_ERR_ARG, "file ends before entry " + std::to_string(i + 1));
			int64_t r, c;
			double v;
			if (!parse_entry(line, r, c, v))
				return fail(FSX_ERR_ARG, "malformed entry line " + std::to_string(i + 1));
			if (cnt >= cap) return fail(FSX_ERR_ARG, "entry capacity exceeded");
			I[cnt] = r;
			J[cnt] = c;
			V[cnt] = v;
			++cnt;
			if (h.symmetric && r != c) {
				if (cnt >= cap) return fail(FSX_ERR_ARG, "entry capacity exceeded");
				I[cnt] = c;
				J[cnt] = r;
				V[cnt] = v;
				++cnt;
			}
		}
		*count = cnt;
		return FSX_OK;
	});
}

int fsx_coo_tocsr(int64_t nrows, int64_t nnz, const int64_t * I, const int64_t * J,
                  const double * V, int64_t * rowptr, int64_t * colind, double * vals) {
	return io_guard([&]() -> int {
		if (nrows < 0 || nnz < 0) return fail(FSX_ERR_ARG, "negative size");
		if (!rowptr || (nnz && (!I || !J || !V || !colind || !vals)))
			return fail(FSX_ERR_ARG, "null argument");
		for (int64_t e = 0; e < nnz; ++e)
			if (I[e] < 0 || I[e] >= nrows) return fail(FSX_ERR_ARG, "row index out of range");
		std::memset(rowptr, 0, (size_t)(nrows + 1) * sizeof(int64_t));
		for (int64_t e = 0; e < nnz; ++e) ++rowptr[I[e] + 1];
		for (int64_t i = 0; i < nrows; ++i) rowptr[i + 1] += rowptr[i];
		std::vector<int64_t> fill(rowptr, rowptr + nrows);
		for (int64_t e = 0; e < nnz; ++e) {
			const int64_t pos = fill[I[e]]++;
			colind[pos] = J[e];
			vals[pos] = V[e];
		}
		return FSX_OK;
	});
}

int fsx_csr_from_mm(fsx_ctx * ctx, const char * path, fsx_itype it, fsx_csr ** out) {
	return io_guard([&]() -> int {
		if (!ctx || !out) return fail(FSX_ERR_ARG, "null argument");
		int64_t nr, nc, lines, cap;
		int sym;
		int rc = fsx_mm_read_header(path, &nr, &nc, &lines, &sym, &cap);
		if (rc) return rc;
		std::vector<int64_t> I(cap), J(cap), rp(nr + 1), ci;
		std::vector<double> V(cap), vv;
		int64_t cnt = 0;
		rc = fsx_mm_read_coo(path, cap, I.data(), J.data(), V.data(), &cnt);
		if (rc) return rc;
		ci.resize(cnt);
		vv.resize(cnt);
		for (int64_t e = 0; e < cnt; ++e)
			if (J[e] < 0 || J[e] >= nc) return fail(FSX_ERR_ARG, "column index out of range");
		rc = fsx_coo_tocsr(nr, cnt, I.data(), J.data(), V.data(), rp.data(), ci.data(), vv.data());
		if (rc) return rc;
		if (it == FSX_I64)
			return fsx_csr_create(ctx, nr, nc, cnt, FSX_I64, rp.data(), ci.data(), FSX_F64,
			                      vv.data(), out);
		if (nr >= ((int64_t)1 << 31) || nc >= ((int64_t)1 << 31) || cnt >= ((int64_t)1 << 31))
			return fail(FSX_ERR_ARG, "int32 indices cannot address this matrix");
		std::vector<int32_t> rp32(rp.begin(), rp.end()), ci32(ci.begin(), ci.end());
		return fsx_csr_create(ctx, nr, nc, cnt, FSX_I32, rp32.data(), ci32.data(), FSX_F64,
		                      vv.data(), out);
	});
}

} // extern "C"
