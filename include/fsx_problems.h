/*
 * fsx_problems.h -- deterministic synthetic inputs for the Krylov path
 * (host code, libfsx_problems.so).  Every generator is a pure function of its
 * arguments and seed; results do not depend on the thread count.
 *
 * Grid convention: unknowns are the interior nodes of an (n+2)^d lattice whose
 * boundary nodes carry Dirichlet u = 0 and are eliminated.  Row id
 * i + n*(j + n*k), x fastest -- the mdcolex / mdspan[k][j][i] order of the
 * reference's structured examples (examples/poisson/poisson.hh:54-66).
 * Columns within a row are ascending.  Coefficient fields live on all
 * (n+2)^d lattice nodes, index I + (n+2)*(J + (n+2)*K) with I = i+1 etc.
 */
#ifndef FSX_PROBLEMS_H
#define FSX_PROBLEMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 2D 5-point [4, -1] Laplacian on an n x n interior grid (nnz 5n^2 - 4n),
 * the Poisson example's stencil with hx = hy (examples/poisson/poisson.cc:61-82,
 * h^2 folded into the right-hand side). */
int64_t fsx_gen_poisson2d_nnz(int n);
int fsx_gen_poisson2d(int n, int32_t * rowptr, int32_t * colind, double * vals);

/* Lognormal coefficient k = exp(sigma * g) on the (n+2)^3 lattice, g a
 * Gaussian random field: N(0,1) white noise from a counter-based generator
 * (splitmix64 of seed and lattice coordinate, Box-Muller), a separable
 * Gaussian filter of standard deviation ell/sqrt(2) cells truncated at 4
 * standard deviations (noise is drawn on a lattice padded by the filter
 * radius, so there are no boundary effects), then shifted and scaled to zero
 * mean and unit variance. */
int fsx_gen_grf_lognormal(int n, double sigma, double ell, uint64_t seed,
                          double * k);

/* 3D 7-point variable-coefficient diffusion: A = shift*I + scale*L_k with
 * L_k u_i = sum_f k_f (u_i - u_nb) over the 6 faces, face coefficient the
 * arithmetic mean k_f = 0.5*(k_hi + k_lo) (the average_coefficient convention,
 * physics/volume_diffusion/coefficient.hh:157-174).  Boundary faces add k_f to
 * the diagonal only.  nnz = 7n^3 - 6n^2.  k is given on the (n+2)^3 lattice. */
int64_t fsx_gen_diffusion3d_nnz(int n);
int fsx_gen_diffusion3d(int n, const double * k, double scale, double shift,
                        int32_t * rowptr, int32_t * colind, double * vals);

/* 3D upwind convection-diffusion, unit diffusion plus first-order upwind
 * convection with cell velocities c = (cx, cy, cz) >= 0:
 * a_ii = 6 + cx + cy + cz, upwind neighbour -1 - c_d, downwind -1. */
int fsx_gen_convdiff3d(int n, double cx, double cy, double cz,
                       int32_t * rowptr, int32_t * colind, double * vals);

/* Batch of independent 3D diffusion systems (block-diagonal concatenation).
 * System s has edge e_s = emin + (splitmix64(seed_e, s) mod (emax-emin+1)),
 * its own lognormal field (fsx_gen_grf_lognormal with seed seed_k + s) and
 * A_s = L_{k_s} as in fsx_gen_diffusion3d.  Rows of system s are
 * [rows[s], rows[s+1]); column ids are global.
 * Call 1: fsx_gen_batch_sizes fills edges[nsys], rows[nsys+1], nnzs[nsys+1].
 * Call 2: fsx_gen_batch fills the concatenated CSR (int32). */
int fsx_gen_batch_sizes(int nsys, int emin, int emax, uint64_t seed_e,
                        int32_t * edges, int64_t * rows, int64_t * nnzs);
int fsx_gen_batch(int nsys, const int32_t * edges, const int64_t * rows,
                  const int64_t * nnzs, double sigma, double ell,
                  uint64_t seed_k, int32_t * rowptr, int32_t * colind,
                  double * vals);

/* x_i ~ U[lo, hi) from splitmix64(seed, i). */
int fsx_gen_uniform(double * x, int64_t n, uint64_t seed, double lo, double hi);

/* Sum of `nb` Gaussian bumps with seeded centres/widths/amplitudes, sampled
 * at interior nodes of the (n+2)^3 lattice with h = 1/(n+1). */
int fsx_gen_bumps(int n, int nb, uint64_t seed, double * u);

#ifdef __cplusplus
}
#endif

#endif
