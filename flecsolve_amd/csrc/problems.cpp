// problems.cpp -- deterministic synthetic inputs (declared in fsx_problems.h).
//
// Host code, OpenMP-parallel; every result is independent of the thread count
// (per-row / per-plane work, reductions combined in a fixed plane order).
#include "fsx_problems.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t z) {
	z += 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

// counter-based stream: value `i` of stream `seed`
inline uint64_t hash2(uint64_t seed, uint64_t i) {
	return splitmix64(splitmix64(seed) ^ (i * 0xD1B54A32D192ED03ull));
}

inline double u01(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }

inline double normal(uint64_t seed, uint64_t i) {
	const double u1 = ((double)(hash2(seed, 2 * i) >> 11) + 1.0) * 0x1.0p-53;
	const double u2 = u01(hash2(seed, 2 * i + 1));
	return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

// Assemble a 7-point (d = 3) or 5-point (d = 2) CSR whose entries come from
// `coef(row, dir) -> value` (dir 0..2d-1 for -z,-y,-x,+x,+y,+z order in 3D and
// -y,-x,+x,+y in 2D) and `diag(row)`.  Neighbours outside the grid are
// dropped.  Columns come out ascending.
template<class Off, class Diag>
void assemble(int d, int n, int32_t * rowptr, int32_t * colind, double * vals,
              Off && off, Diag && diag) {
	const int64_t N = d == 3 ? (int64_t)n * n * n : (int64_t)n * n;
	// per-row counts
	std::vector<int32_t> cnt(N);
#pragma omp parallel for schedule(static)
	for (int64_t r = 0; r < N; ++r) {
		const int i = (int)(r % n);
		const int j = (int)((r / n) % n);
		const int k = d == 3 ? (int)(r / ((int64_t)n * n)) : 0;
		int c = 1 + (i > 0) + (i < n - 1) + (j > 0) + (j < n - 1);
		if (d == 3) c += (k > 0) + (k < n - 1);
		cnt[r] = c;
	}
	rowptr[0] = 0;
	for (int64_t r = 0; r < N; ++r) rowptr[r + 1] = rowptr[r] + cnt[r];
#pragma omp parallel for schedule(static)
	for (int64_t r = 0; r < N; ++r) {
		const int i = (int)(r % n);
		const int j = (int)((r / n) % n);
		const int k = d == 3 ? (int)(r / ((int64_t)n * n)) : 0;
		const int64_t nn = (int64_t)n * n;
		int64_t e = rowptr[r];
		auto put = [&](int64_t c, double v) {
			colind[e] = (int32_t)c;
			vals[e] = v;
			++e;
		};
		if (d == 3) {
			if (k > 0) put(r - nn, off(r, 0));
			if (j > 0) put(r - n, off(r, 1));
			if (i > 0) put(r - 1, off(r, 2));
			put(r, diag(r));
			if (i < n - 1) put(r + 1, off(r, 3));
			if (j < n - 1) put(r + n, off(r, 4));
			if (k < n - 1) put(r + nn, off(r, 5));
		}
		else {
			if (j > 0) put(r - n, off(r, 0));
			if (i > 0) put(r - 1, off(r, 1));
			put(r, diag(r));
			if (i < n - 1) put(r + 1, off(r, 2));
			if (j < n - 1) put(r + n, off(r, 3));
		}
	}
}

} // namespace

extern "C" {

int64_t fsx_gen_poisson2d_nnz(int n) { return 5ll * n * n - 4ll * n; }

int fsx_gen_poisson2d(int n, int32_t * rowptr, int32_t * colind,
                      double * vals) {
	if (n < 1) return 1;
	assemble(
		2, n, rowptr, colind, vals, [](int64_t, int) { return -1.0; },
		[](int64_t) { return 4.0; });
	return 0;
}

int fsx_gen_grf_lognormal(int n, double sigma, double ell, uint64_t seed,
                          double * k) {
	if (n < 1 || ell <= 0) return 1;
	const double s = ell / std::sqrt(2.0);
	const int R = (int)std::ceil(4.0 * s);
	std::vector<double> w(2 * R + 1);
	double wsum = 0;
	for (int m = -R; m <= R; ++m) {
		w[m + R] = std::exp(-(double)m * m / (2.0 * s * s));
		wsum += w[m + R];
	}
	for (auto & v : w) v /= wsum;
	const int L = n + 2;          // lattice nodes per axis
	const int M = L + 2 * R;      // padded
	const int64_t M2 = (int64_t)M * M;
	std::vector<double> noise((size_t)M * M * M);
#pragma omp parallel for schedule(static)
	for (int64_t idx = 0; idx < (int64_t)M * M2; ++idx)
		noise[idx] = normal(seed, (uint64_t)idx);
	// filter along x: [M][M][L]
	std::vector<double> t1((size_t)M * M * L);
#pragma omp parallel for schedule(static)
	for (int64_t kj = 0; kj < M2; ++kj) {
		const double * src = &noise[kj * M];
		double * dst = &t1[kj * L];
		for (int i = 0; i < L; ++i) {
			double a = 0;
			for (int m = 0; m <= 2 * R; ++m) a += w[m] * src[i + m];
			dst[i] = a;
		}
	}
	std::vector<double>().swap(noise);
	// filter along y: [M][L][L]
	std::vector<double> t2((size_t)M * L * L);
#pragma omp parallel for schedule(static)
	for (int64_t kk = 0; kk < M; ++kk)
		for (int j = 0; j < L; ++j) {
			double * dst = &t2[((size_t)kk * L + j) * L];
			for (int i = 0; i < L; ++i) dst[i] = 0;
			for (int m = 0; m <= 2 * R; ++m) {
				const double * src = &t1[((size_t)kk * M + (j + m)) * L];
				const double wm = w[m];
				for (int i = 0; i < L; ++i) dst[i] += wm * src[i];
			}
		}
	std::vector<double>().swap(t1);
	// filter along z: [L][L][L] into k (as g for now)
	const int64_t L2 = (int64_t)L * L;
#pragma omp parallel for schedule(static)
	for (int64_t kk = 0; kk < L; ++kk) {
		double * dst = &k[kk * L2];
		for (int64_t q = 0; q < L2; ++q) dst[q] = 0;
		for (int m = 0; m <= 2 * R; ++m) {
			const double * src = &t2[(size_t)(kk + m) * L2];
			const double wm = w[m];
			for (int64_t q = 0; q < L2; ++q) dst[q] += wm * src[q];
		}
	}
	std::vector<double>().swap(t2);
	// zero mean, unit variance (per-plane partials combined in plane order)
	std::vector<double> ps(L), pq(L);
#pragma omp parallel for schedule(static)
	for (int64_t kk = 0; kk < L; ++kk) {
		double a = 0;
		for (int64_t q = 0; q < L2; ++q) a += k[kk * L2 + q];
		ps[kk] = a;
	}
	double sum = 0;
	for (int kk = 0; kk < L; ++kk) sum += ps[kk];
	const double N = (double)L * L2;
	const double mean = sum / N;
#pragma omp parallel for schedule(static)
	for (int64_t kk = 0; kk < L; ++kk) {
		double a = 0;
		for (int64_t q = 0; q < L2; ++q) {
			const double d = k[kk * L2 + q] - mean;
			a += d * d;
		}
		pq[kk] = a;
	}
	double var = 0;
	for (int kk = 0; kk < L; ++kk) var += pq[kk];
	const double sd = std::sqrt(var / N);
	const double inv = sd > 0 ? 1.0 / sd : 0.0;
#pragma omp parallel for schedule(static)
	for (int64_t q = 0; q < (int64_t)L * L2; ++q)
		k[q] = std::exp(sigma * ((k[q] - mean) * inv));
	return 0;
}

int64_t fsx_gen_diffusion3d_nnz(int n) {
	return 7ll * n * n * n - 6ll * n * n;
}

int fsx_gen_diffusion3d(int n, const double * k, double scale, double shift,
                        int32_t * rowptr, int32_t * colind, double * vals) {
	if (n < 1) return 1;
	const int64_t L = n + 2, L2 = L * L;
	auto node = [&](int64_t r) {
		const int64_t i = r % n, j = (r / n) % n, kk = r / ((int64_t)n * n);
		return (i + 1) + L * ((j + 1) + L * (kk + 1));
	};
	// lattice offsets for -z,-y,-x,+x,+y,+z
	const int64_t noff[6] = {-L2, -L, -1, 1, L, L2};
	// face coefficient between lattice node a and its neighbour in dir d;
	// written as 0.5 * (k_hi + k_lo) with hi the node on the + side.
	auto kface = [&](int64_t a, int d) {
		const int64_t b = a + noff[d];
		return d < 3 ? 0.5 * (k[a] + k[b]) : 0.5 * (k[b] + k[a]);
	};
	assemble(
		3, n, rowptr, colind, vals,
		[&](int64_t r, int d) { return -scale * kface(node(r), d); },
		[&](int64_t r) {
			const int64_t a = node(r);
			double s = 0;
			for (int d = 0; d < 6; ++d) s += kface(a, d);
			return shift + scale * s;
		});
	return 0;
}

int fsx_gen_convdiff3d(int n, double cx, double cy, double cz,
                       int32_t * rowptr, int32_t * colind, double * vals) {
	if (n < 1 || cx < 0 || cy < 0 || cz < 0) return 1;
	const double c[3] = {cz, cy, cx}; // dir order -z,-y,-x
	assemble(
		3, n, rowptr, colind, vals,
		[&](int64_t, int d) {
			// d 0..2: upwind (-z,-y,-x); d 3..5: downwind (+x,+y,+z)
			return d < 3 ? -1.0 - c[d] : -1.0;
		},
		[&](int64_t) { return 6.0 + cx + cy + cz; });
	return 0;
}

int fsx_gen_batch_sizes(int nsys, int emin, int emax, uint64_t seed_e,
                        int32_t * edges, int64_t * rows, int64_t * nnzs) {
	if (nsys < 0 || emin < 1 || emax < emin) return 1;
	rows[0] = 0;
	nnzs[0] = 0;
	for (int s = 0; s < nsys; ++s) {
		const uint64_t span = (uint64_t)(emax - emin + 1);
		const int e = emin + (int)(hash2(seed_e, (uint64_t)s) % span);
		edges[s] = e;
		rows[s + 1] = rows[s] + (int64_t)e * e * e;
		nnzs[s + 1] = nnzs[s] + fsx_gen_diffusion3d_nnz(e);
	}
	return 0;
}

int fsx_gen_batch(int nsys, const int32_t * edges, const int64_t * rows,
                  const int64_t * nnzs, double sigma, double ell,
                  uint64_t seed_k, int32_t * rowptr, int32_t * colind,
                  double * vals) {
	if (nsys < 0) return 1;
	if (nnzs[nsys] >= ((int64_t)1 << 31)) return 2;
	int rc = 0;
	// one system per iteration; inner generators run single-threaded here
#pragma omp parallel for schedule(dynamic, 1) reduction(| : rc)
	for (int s = 0; s < nsys; ++s) {
		const int e = edges[s];
		const int64_t L = e + 2;
		std::vector<double> k((size_t)(L * L * L));
		std::vector<int32_t> rp(rows[s + 1] - rows[s] + 1);
		int r1 = fsx_gen_grf_lognormal(e, sigma, ell, seed_k + (uint64_t)s, k.data());
		int r2 = fsx_gen_diffusion3d(e, k.data(), 1.0, 0.0, rp.data(),
		                             colind + nnzs[s], vals + nnzs[s]);
		rc |= r1 | r2;
		for (int64_t i = 0; i < (int64_t)rp.size() - 1; ++i)
			rowptr[rows[s] + i] = (int32_t)(nnzs[s] + rp[i]);
		for (int64_t q = nnzs[s]; q < nnzs[s + 1]; ++q)
			colind[q] += (int32_t)rows[s];
	}
	rowptr[rows[nsys]] = (int32_t)nnzs[nsys];
	return rc;
}

int fsx_gen_uniform(double * x, int64_t n, uint64_t seed, double lo,
                    double hi) {
#pragma omp parallel for schedule(static)
	for (int64_t i = 0; i < n; ++i)
		x[i] = lo + (hi - lo) * u01(hash2(seed, (uint64_t)i));
	return 0;
}

int fsx_gen_bumps(int n, int nb, uint64_t seed, double * u) {
	if (n < 1 || nb < 0) return 1;
	std::vector<double> cx(nb), cy(nb), cz(nb), wd(nb), am(nb);
	for (int b = 0; b < nb; ++b) {
		cx[b] = 0.2 + 0.6 * u01(hash2(seed, 5 * b + 0));
		cy[b] = 0.2 + 0.6 * u01(hash2(seed, 5 * b + 1));
		cz[b] = 0.2 + 0.6 * u01(hash2(seed, 5 * b + 2));
		wd[b] = 0.05 + 0.1 * u01(hash2(seed, 5 * b + 3));
		am[b] = 0.5 + 1.0 * u01(hash2(seed, 5 * b + 4));
	}
	const double h = 1.0 / (n + 1);
	const int64_t N = (int64_t)n * n * n;
#pragma omp parallel for schedule(static)
	for (int64_t r = 0; r < N; ++r) {
		const double x = (double)(r % n + 1) * h;
		const double y = (double)((r / n) % n + 1) * h;
		const double z = (double)(r / ((int64_t)n * n) + 1) * h;
		double s = 0;
		for (int b = 0; b < nb; ++b) {
			const double d2 = (x - cx[b]) * (x - cx[b]) + (y - cy[b]) * (y - cy[b]) +
			                  (z - cz[b]) * (z - cz[b]);
			s += am[b] * std::exp(-d2 / (2 * wd[b] * wd[b]));
		}
		u[r] = s;
	}
	return 0;
}

} // extern "C"
