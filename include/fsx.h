/*
 * fsx.h -- C ABI of libfsx, the MI355X (gfx950) implementation of flecsolve's
 * Krylov solve path: device vectors, a device CSR operator, fused CG kernels
 * and solve drivers.
 *
 * Conventions
 *   - Every call returns an int status: 0 = OK, non-zero = error; the message
 *     is available from fsx_last_error() (thread-local).  No C++ exception
 *     crosses this boundary.  Contract violations that the reference turns
 *     into flog_assert aborts return FSX_ERR_ARG here.
 *   - Handles are opaque.  One context per host thread; each context owns a
 *     HIP stream (default priority) and all calls on it are stream-ordered.
 *     Calls that return host values (reductions into double*, downloads,
 *     solve drivers) synchronise the context stream before returning.
 *   - Vectors are reference-counted: copying a handle (fsx_vec_retain) makes
 *     an alias of the same device buffer, which is the ownership model of the
 *     reference's topo_view vectors (vectors/data/topo_view.hh:73-81, equality
 *     by field id).  fsx_vec_id() returns the identity used for that equality.
 *   - Elementwise operations are alias-safe: any output may be the same
 *     handle as any input (the reference dispatches aliasing cases to *_self
 *     tasks, vectors/operations/topo_view.hh:78-210).  fsx_scale_to() rejects
 *     aliasing, as the reference asserts (topo_view.hh:71-76).
 */
#ifndef FSX_H
#define FSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
	FSX_OK = 0,
	FSX_ERR_ARG = 1,     /* invalid argument / contract violation */
	FSX_ERR_HIP = 2,     /* HIP runtime error */
	FSX_ERR_ALLOC = 3,   /* device or host allocation failure */
	FSX_ERR_STATE = 4    /* object in the wrong state */
};

typedef enum { FSX_F64 = 0, FSX_F32 = 1 } fsx_dtype;
typedef enum { FSX_I32 = 0, FSX_I64 = 1 } fsx_itype;

typedef struct fsx_ctx fsx_ctx;
typedef struct fsx_vec fsx_vec;
typedef struct fsx_csr fsx_csr;

/* solver_settings (flecsolve/solvers/solver_settings.hh:31-36) plus the
 * GMRES settings (flecsolve/solvers/gmres.hh:389-393).  rtol/atol are float
 * exactly as in the reference. */
typedef struct {
	int maxiter;
	float rtol;
	float atol;
	int use_zero_guess;
	int max_krylov_dim; /* GMRES: < 0 means maxiter */
	int pre_side;       /* GMRES: 0 = left, 1 = right (reference default) */
	int restart;        /* GMRES: non-zero enables restarts */
} fsx_settings;

/* solve_info (flecsolve/solvers/solver_settings.hh:61-85); norms are float
 * as in the reference.  status follows solve_info::stop_reason order. */
typedef struct {
	int status;
	int iters;
	int restarts;
	float res_norm_initial, res_norm_final;
	float sol_norm_initial, sol_norm_final;
	float rhs_norm;
} fsx_info;

enum {
	FSX_CONVERGED_ATOL = 0,
	FSX_CONVERGED_RTOL = 1,
	FSX_CONVERGED_USER = 2,
	FSX_DIVERGED_DTOL = 3,
	FSX_DIVERGED_ITERS = 4,
	FSX_DIVERGED_BREAKDOWN = 5,
	FSX_UNKNOWN = 6
};

/* ---- context ----------------------------------------------------------
 * Replaces the FleCSI runtime/scheduler the reference's tasks run on
 * (flecsi::scheduler::instance, vectors/operations/topo_view.hh:38-40). */
const char * fsx_last_error(void);
int fsx_ctx_create(int device, fsx_ctx ** out);
int fsx_ctx_destroy(fsx_ctx * ctx);
int fsx_sync(fsx_ctx * ctx);
/* The context's hipStream_t, for callers that time or order their own work
 * against the library's kernels. */
int fsx_ctx_stream(fsx_ctx * ctx, void ** stream);
/* Per-context counters of device reductions forced to the host (each is one
 * blocking .get() of the reference, topo_view.hh:265-269). */
int fsx_ctx_counters(fsx_ctx * ctx, uint64_t * host_syncs, uint64_t * launches);
/* Device allocations the library has made in this process (hipMalloc calls,
 * all contexts) and the bytes its live allocations hold now.  No counterpart
 * in the reference: lets a caller confirm that an entry point allocated
 * nothing. */
int fsx_alloc_stats(uint64_t * allocations, uint64_t * live_bytes);

/* ---- vectors: vec::core data (vectors/core.hh:26-353) ------------------ */
int fsx_vec_create(fsx_ctx * ctx, size_t n, fsx_dtype dt, fsx_vec ** out);
/* Non-owning view of existing device memory (e.g. a torch tensor). */
int fsx_vec_wrap(fsx_ctx * ctx, void * dptr, size_t n, fsx_dtype dt,
                 fsx_vec ** out);
int fsx_vec_retain(fsx_vec * v);
int fsx_vec_release(fsx_vec * v);
int fsx_vec_size(const fsx_vec * v, size_t * n);
int fsx_vec_dtype(const fsx_vec * v, int * dt);
int fsx_vec_id(const fsx_vec * v, uint64_t * id);
int fsx_vec_device_ptr(const fsx_vec * v, void ** dptr);
int fsx_vec_upload(fsx_ctx * ctx, fsx_vec * v, const void * host, size_t n);
int fsx_vec_download(fsx_ctx * ctx, const fsx_vec * v, void * host, size_t n);
/* One element, for vec::core::operator[] / Ops::retreive (vectors/core.hh:
 * 338-340, vectors/seq.hh:271-280): read (converted to double) or write
 * element i, ordered after the context's queued work; both synchronise. */
int fsx_vec_get(fsx_ctx * ctx, const fsx_vec * v, size_t i, double * out);
int fsx_vec_set(fsx_ctx * ctx, fsx_vec * v, size_t i, double value);

/* ---- BLAS-1, mirroring the Ops policy (vectors/core.hh:59-310;
 *      semantics vectors/seq.hh:93-304; tasks operations/topo_tasks.hh) ---- */
int fsx_copy(fsx_ctx * ctx, const fsx_vec * src, fsx_vec * dst);          /* core.hh:59-62 */
int fsx_zero(fsx_ctx * ctx, fsx_vec * x);                                  /* core.hh:67 */
int fsx_set_scalar(fsx_ctx * ctx, double a, fsx_vec * x);                 /* core.hh:75 */
int fsx_scale(fsx_ctx * ctx, double a, fsx_vec * x);                       /* core.hh:95 */
int fsx_scale_to(fsx_ctx * ctx, double a, const fsx_vec * x, fsx_vec * y);/* core.hh:84-87 */
int fsx_add(fsx_ctx * ctx, const fsx_vec * x, const fsx_vec * y, fsx_vec * z);      /* core.hh:104-110 */
int fsx_subtract(fsx_ctx * ctx, const fsx_vec * x, const fsx_vec * y, fsx_vec * z); /* core.hh:119-125 */
int fsx_multiply(fsx_ctx * ctx, const fsx_vec * x, const fsx_vec * y, fsx_vec * z); /* core.hh:134-140 */
int fsx_divide(fsx_ctx * ctx, const fsx_vec * x, const fsx_vec * y, fsx_vec * z);   /* core.hh:149-155 */
int fsx_reciprocal(fsx_ctx * ctx, const fsx_vec * x, fsx_vec * y);        /* core.hh:162-165 */
int fsx_linear_sum(fsx_ctx * ctx, double a, const fsx_vec * x, double b,
                   const fsx_vec * y, fsx_vec * z);                        /* core.hh:172-178 */
int fsx_axpy(fsx_ctx * ctx, double a, const fsx_vec * x, const fsx_vec * y,
             fsx_vec * z);                                                 /* core.hh:185-191 */
int fsx_axpby(fsx_ctx * ctx, double a, double b, const fsx_vec * x,
              fsx_vec * z);                                                /* core.hh:198-201 */
int fsx_abs(fsx_ctx * ctx, const fsx_vec * x, fsx_vec * y);               /* core.hh:208-211 */
int fsx_add_scalar(fsx_ctx * ctx, const fsx_vec * x, double a, fsx_vec * y); /* core.hh:218-221 */
/* mt19937(seed) + U[0,1) doubles, as seq_ops::set_random (seq.hh:103-115) */
int fsx_set_random(fsx_ctx * ctx, fsx_vec * x, unsigned seed);            /* core.hh:302-310 */
/* FP64 <-> FP32 conversion (new: mixed-precision refinement) */
int fsx_convert(fsx_ctx * ctx, const fsx_vec * src, fsx_vec * dst, double scale);

/* ---- reductions (vectors/core.hh:229-290; topo_tasks.hh:52-66, 216-303).
 * All accumulate in FP64 with a fixed, size-determined order (deterministic
 * run to run).  Each returns the value to the host (synchronises). ------ */
int fsx_dot(fsx_ctx * ctx, const fsx_vec * x, const fsx_vec * y, double * out);
/* Device-scalar flavour of the reductions: the result goes to device memory
 * at d_out (8 bytes), stream-ordered, no host synchronisation (the value a
 * reference future's .get() would return; FSX_RED_L2SQ is the squared
 * 2-norm). */
enum { FSX_RED_DOT = 0, FSX_RED_L1 = 1, FSX_RED_L2SQ = 2, FSX_RED_INF = 3,
       FSX_RED_MIN = 4, FSX_RED_MAX = 5 };
int fsx_reduce_async(fsx_ctx * ctx, int kind, const fsx_vec * x, const fsx_vec * y,
                     double * d_out);
int fsx_nrm1(fsx_ctx * ctx, const fsx_vec * x, double * out);
int fsx_nrm2(fsx_ctx * ctx, const fsx_vec * x, double * out);
int fsx_nrminf(fsx_ctx * ctx, const fsx_vec * x, double * out);
/* lp_norm<p> of the FleCSI vector path (vectors/operations/topo_view.hh:
 * 243-257, topo_tasks.hh:216-226): p = 1 and 2 as fsx_nrm1 / fsx_nrm2;
 * p >= 3 gives (sum_i pow(x_i, p))^(1/p) -- with no absolute value, as
 * there (NaN when an odd p makes the sum negative).  The sum is folded in
 * the vector's real type, as lp_norm_local folds into `real`: double for
 * FP64 vectors, float for FP32 ones (each term rounded to float). */
int fsx_nrmp(fsx_ctx * ctx, const fsx_vec * x, int p, double * out);
int fsx_min(fsx_ctx * ctx, const fsx_vec * x, double * out);
int fsx_max(fsx_ctx * ctx, const fsx_vec * x, double * out);

/* ---- CSR operator: replaces mat::parcsr / csr apply()
 *      (matrices/parcsr.hh:60-91 -> matrices/seq.hh:177-193) ------------- */
int fsx_csr_create(fsx_ctx * ctx, int64_t nrows, int64_t ncols, int64_t nnz,
                   fsx_itype it, const void * rowptr, const void * colind,
                   fsx_dtype vt, const void * vals, fsx_csr ** out);
/* New matrix sharing A's row/column index arrays, with values converted to
 * `vt` (FP32 values for mixed-precision refinement). */
int fsx_csr_convert(fsx_ctx * ctx, const fsx_csr * A, fsx_dtype vt,
                    fsx_csr ** out);
int fsx_csr_destroy(fsx_csr * A);
int fsx_csr_info(const fsx_csr * A, int64_t * nrows, int64_t * ncols,
                 int64_t * nnz, int * itype, int * vtype);
/* Row-mapping selected for A: 0 = LDS-staged row blocks, 1 = sub-wavefront
 * vector per row (`lanes` is the vector width), 2 = sliced rows (64-row
 * slices, entry-major, column per entry), 3 = slice offsets (64-row slices
 * storing the distinct column offsets of their rows and lane masks instead
 * of a column per entry; chosen for stencil-like structure), 4 = symmetric
 * stencil (one offset list of at most 8 offsets for the whole matrix, lane
 * masks, values stored only for the offsets >= 0; the entries below the
 * diagonal are read from their mirrors above it, so it serves only
 * structurally and bitwise symmetric values; chosen for such FP64 matrices,
 * with slice offsets kept beside it).  A change of values that breaks the
 * symmetry moves a plan-4 matrix to plan 3; fsx_csr_set_plan(A, 4, 0)
 * fails for a structure or values that do not fit.  Every mapping sums each
 * row serially in CSR order except 1. */
int fsx_csr_plan(const fsx_csr * A, int * kind, int * lanes);
int fsx_csr_set_plan(fsx_csr * A, int kind, int lanes);
/* Replace A's values (nnz entries of A's value type, CSR order) from host
 * memory; the SpMV layouts are refreshed in place (captured solver graphs
 * stay valid). */
int fsx_csr_set_values(fsx_ctx * ctx, fsx_csr * A, const void * vals);
/* A32 (= fsx_csr_convert(A64, FSX_F32), shared structure) takes A64's
 * current values rounded to FP32, on the device. */
int fsx_csr_set_values_f32_from_f64(fsx_ctx * ctx, fsx_csr * A32, const fsx_csr * A64);
/* A32 (= fsx_csr_convert(A64, FSX_F32)) takes S A64 S rounded to FP32,
 * S = diag(1 / sqrt(a_ii)) in FP64: a32_ij = (float)((s_i a_ij) s_j).  Every
 * diagonal entry of A64 must be stored and positive.  The inner matrix of
 * fsx_ir_cg_solve with jacobi = 2.  New, no reference counterpart. */
int fsx_csr_set_values_f32_sym_scaled(fsx_ctx * ctx, fsx_csr * A32, const fsx_csr * A64);
/* A = alpha I + beta F on F's pattern (A = fsx_csr_convert(F, FSX_F64), or
 * any matrix sharing F's index arrays; FP64, square, every row storing its
 * diagonal entry unless alpha == 0): each entry beta f_ij, the row's
 * diagonal entry (the last one if the row stores several) alpha + beta f_ii,
 * computed on the device; A's SpMV layouts are rewritten in place.  Forms the implicit stage matrix I - gamma F of
 * time_integrator::operator_adapter (time-integrators/operator_adapter.hh:
 * 29-35, y = x - gamma f(x)) as one matrix. */
int fsx_csr_shift_scale(fsx_ctx * ctx, const fsx_csr * F, double alpha, double beta,
                        fsx_csr * A);
int fsx_csr_download(fsx_ctx * ctx, const fsx_csr * A, void * rowptr,
                     void * colind, void * vals);
/* y = A x   (sparse::apply -> mult -> Ops::spmv, matrices/seq.hh:46-60) */
int fsx_csr_spmv(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * x,
                 fsx_vec * y);
/* r = b - A x  (op::core::residual, operators/core.hh:129-142), fused */
int fsx_csr_residual(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * b,
                     const fsx_vec * x, fsx_vec * r);
/* dinv_i = 1 / a_ii (values of the Dinv CSR of util/test/mesh.hh:123-140) */
int fsx_csr_diag_inv(fsx_ctx * ctx, const fsx_csr * A, fsx_vec * dinv);
/* y = A^T x (adjoint solves) */
int fsx_csr_spmv_t(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * x,
                   fsx_vec * y);

/* ---- parallel CSR: topo::csr colouring + mat::parcsr apply, all colours on
 *      one device --------------------------------------------------------
 * fsx_parcsr_create splits the FP64 matrix A over `ncolours` colours as
 * topo::csr::color / init_mats do (topo/csr.hh:482-618): FleCSI equal_map
 * row and column blocks; each owned row's entries, in CSR order, go to the
 * colour's diag block (column in the colour's column block, local id) or
 * its offd block (column ncols_local + rank in the colour's sorted unique
 * ghost list, colmap).  A is copied (it may be destroyed afterwards).
 * ncolours in [1, min(nrows, ncols)]. */
typedef struct fsx_parcsr fsx_parcsr;
int fsx_parcsr_create(fsx_ctx * ctx, const fsx_csr * A, int ncolours, fsx_parcsr ** out);
int fsx_parcsr_destroy(fsx_parcsr * M);
/* Colour c's index-space sizes (color(), csr.hh:524-528): sizes[0] rows,
 * [1] local columns, [2] ghost columns, [3] diag entries, [4] offd entries. */
int fsx_parcsr_colour_info(const fsx_parcsr * M, int colour, int64_t * sizes);
/* Colour c's blocks as the reference stores them (int64 indices): diag
 * rowptr (rows + 1) / colind (local ids) / vals, offd rowptr / colind
 * (ncols_local + ghost rank) / vals, colmap (ghost global ids).  Any
 * output may be NULL. */
int fsx_parcsr_download(fsx_ctx * ctx, const fsx_parcsr * M, int colour,
                        int64_t * diag_rowptr, int64_t * diag_colind, double * diag_vals,
                        int64_t * offd_rowptr, int64_t * offd_colind, double * offd_vals,
                        int64_t * colmap);
/* y = A x as parcsr_ops::spmv computes it over the colours
 * (matrices/parcsr.hh:60-91): ghost copy of x, tmp = offd x_ghost,
 * y = diag x_local, y = y + tmp, each block summed from 0 in CSR order.
 * Bit-identical to the reference's ncolours-colour product.  ctx must be
 * the context M was created on (M's ghost buffer is ordered on its
 * stream). */
int fsx_parcsr_spmv(fsx_ctx * ctx, const fsx_parcsr * M, const fsx_vec * x, fsx_vec * y);
/* r = b - A x (op::core::residual, operators/core.hh:129-142). */
int fsx_parcsr_residual(fsx_ctx * ctx, const fsx_parcsr * M, const fsx_vec * b,
                        const fsx_vec * x, fsx_vec * r);
/* The diag blocks of all colours as one CSR with global column ids (owned
 * by M). */
int fsx_parcsr_diag(const fsx_parcsr * M, fsx_csr ** D);

/* ---- fused kernels (CG iteration pieces, cg.hh:95-130) ----------------- */
/* w = A p and out = w . p in one pass (cg.hh:95-98) */
int fsx_spmv_dot(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * p,
                 fsx_vec * w, double * out);
/* z = dinv .* r  (Jacobi preconditioner apply; Dinv SpMV of mesh.hh:123-140) */
int fsx_jacobi_apply(fsx_ctx * ctx, const fsx_vec * dinv, const fsx_vec * r,
                     fsx_vec * z);
/* Weighted Jacobi relaxation, nrelax sweeps, omega stored as float
 * (mg::bound_jacobi::relax, solvers/mg/jacobi.hh:58-93). */
int fsx_jacobi_relax(fsx_ctx * ctx, const fsx_csr * A, float omega,
                     int nrelax, const fsx_vec * b, fsx_vec * x);

/* ---- hybrid Gauss-Seidel smoother ----------------------------------------
 * op::hybrid_gs (solvers/mg/gs.hh:40-130): Gauss-Seidel inside each of
 * `ncolours` row blocks (FleCSI equal_map partition), the other blocks'
 * values taken from before the sweep; ncolours = 1 is the sequential
 * sweep of the serial backend.  Bit-identical to the sequential
 * reference.  fsx_gs_create analyses the structure of the square FP64
 * matrix A (which must outlive the smoother; its values may change in
 * place).  sweep: 0 forward, 1 backward, 2 symmetric (mg::relax_sweep);
 * x is relaxed in place `nrelax` times (hybrid_gs::apply, gs.hh:56-76). */
typedef struct fsx_gs fsx_gs;
int fsx_gs_create(fsx_ctx * ctx, const fsx_csr * A, int ncolours, fsx_gs ** out);
int fsx_gs_relax(fsx_ctx * ctx, fsx_gs * G, int sweep, int nrelax, const fsx_vec * b,
                 fsx_vec * x);
/* Launch levels per forward / backward sweep (the dependency depth). */
int fsx_gs_levels(const fsx_gs * G, int64_t * forward, int64_t * backward);
int fsx_gs_destroy(fsx_gs * G);

/* ---- unsmoothed-aggregation AMG (mg::ua::bound_solver, one colour) -------
 * Setup (solvers/mg/ua.hh:110-139; coarsen.hh:26-655): double (or triple)
 * pairwise aggregation with the reference's dominance check and priority
 * order, Galerkin coarse operators, levels until one has fewer than
 * min_coarse rows or max_levels exist, dense LU of the coarsest (at most
 * 4096 rows).  Apply (ua.hh:141-172): x = 0, then up to maxiter cycles
 * (cycle.hh: 0 v, 1 relaxed-w, 2 relaxed-kappa, 3 kappa-k) with nrelax
 * symmetric hybrid Gauss-Seidel sweeps before and after each coarse
 * correction, stopping once |b - A x| <= atol.  V, relaxed W and relaxed
 * kappa cycles are bit-identical to the sequential algorithm.  A must
 * outlive the solver.  Reference option defaults: min_coarse 5, maxiter
 * 10, atol 1e-6, beta 0.25, pairwise_passes 2, nrelax 2, cycle v,
 * kappa 1, ktol 0.25 (ua.cc). */
typedef struct {
	int max_levels;
	int64_t min_coarse;
	int maxiter;
	float atol;
	float beta;
	int pairwise_passes;
	int nrelax;
	int cycle;
	int kappa;
	float ktol;
	int smoother;        /* 0 hybrid Gauss-Seidel (create_smoother, ua.hh:184-197),
	                        1 weighted Jacobi with jacobi_weight (the alternative
	                        create_smoother keeps commented out, ua.hh:188-190) */
	float jacobi_weight; /* 0.6666666 */
} fsx_amg_settings;
typedef struct fsx_amg fsx_amg;
int fsx_amg_create(fsx_ctx * ctx, const fsx_csr * A, const fsx_amg_settings * s,
                   fsx_amg ** out);
/* Number of levels; rows and stored entries of the first `cap` levels. */
int fsx_amg_levels(const fsx_amg * M, int * nlevels, int64_t * rows, int64_t * nnz, int cap);
/* Aggregate of each row of `level` on level + 1 (-1: in no aggregate). */
int fsx_amg_aggregates(const fsx_amg * M, int level, int64_t * aggt);
/* The matrix of `level` (owned by the solver; level 0 is A). */
int fsx_amg_level_matrix(const fsx_amg * M, int level, fsx_csr ** A);
/* x = the solver applied to b; cycles run and the final |b - A x|. */
int fsx_amg_apply(fsx_ctx * ctx, fsx_amg * M, const fsx_vec * b, fsx_vec * x, int * cycles,
                  double * res);
int fsx_amg_destroy(fsx_amg * M);
/* CG (flexible = 0) or FCG (flexible = 1) preconditioned by M: the generic
 * op::cg / op::fcg (solvers/cg.hh) with mg::ua::bound_solver as P, as a
 * caller of the reference writes cg::solver(s, work)(op::ref(A),
 * op::ref(amg)).  M's fine matrix must be A. */
int fsx_krylov_cg_amg(fsx_ctx * ctx, const fsx_csr * A, fsx_amg * M, int flexible,
                      const fsx_vec * b, fsx_vec * x, const fsx_settings * s, fsx_info * info,
                      double * res_hist, int hist_cap);

/* ---- solve drivers -------------------------------------------------------
 * fsx_cg_solve: op::cg::apply (solvers/cg.hh:43-139) with identity (dinv ==
 * NULL) or Jacobi preconditioning, run as fused kernels with device-resident
 * scalars and device-side convergence flags, captured in HIP graphs of
 * `chunk` iterations.  Same iteration sequence, tolerance semantics and
 * solve_info as the reference (no user diagnostic). */
int fsx_cg_solve(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv,
                 const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                 fsx_info * info);
/* fsx_cg_solve with the reference's user diagnostic (cg.hh:112-116,
 * krylov_parameters.hh:29-31): after each iteration's x and r updates the
 * library calls diag(user, x, |r|) on the host thread, before the tolerance
 * test; a positive return stops the solve with FSX_CONVERGED_USER and
 * iters = iteration + 1, zero continues, a negative one aborts the solve
 * with FSX_ERR_STATE.  The callback may run library operations on x (same
 * context).  Same kernels as fsx_cg_solve, one iteration at a time (one host
 * synchronisation per iteration instead of the reference's three). */
typedef int (*fsx_diag_fn)(void * user, fsx_vec * x, double res_norm);
int fsx_cg_solve_diag(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv,
                      const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                      fsx_diag_fn diag, void * user, fsx_info * info);
/* fsx_cg_solve_noalloc: no counterpart in the reference.  fsx_cg_solve for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same kernels,
 * iterates, iteration count and solve_info as fsx_cg_solve; it allocates
 * nothing: the iteration state, setup scalars, reduction partials and the
 * work vectors r, z, p, w live in a static device array of each GPU (34 MiB,
 * held from the library's load until the process ends), whose address
 * fsx_ctx_create looks up on its device.  Iterations are launched directly
 * (no graphs), the host reading the state after each chunk; no residual
 * history is recorded (fsx_cg_history keeps the last fsx_cg_solve's).
 * Limit: n <= FSX_NOALLOC_MAX_ROWS (the reduction partials of every SpMV
 * plan then fit too); larger systems return FSX_ERR_ARG.  All calls on one
 * GPU share the array and it keeps the last call's contents (every solve
 * writes what it reads first): calls that overlap on one GPU (several host
 * threads or contexts) race. */
#define FSX_NOALLOC_MAX_ROWS ((size_t)1 << 20)
int fsx_cg_solve_noalloc(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv,
                         const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                         fsx_info * info);
/* Residual history of the last fsx_cg_solve on ctx (|r| per iteration). */
int fsx_cg_history(fsx_ctx * ctx, double * hist, int cap, int * count);
/* One pass of cg.hh:108-127: x += alpha p; r -= alpha w; z = D^{-1} r (when
 * dinv is given); rr = r.r and rz = r.z (r.r without dinv). */
int fsx_cg_update(fsx_ctx * ctx, double alpha, const fsx_vec * p, const fsx_vec * w,
                  const fsx_vec * dinv, fsx_vec * x, fsx_vec * r, fsx_vec * z, double * rr,
                  double * rz);

/* fsx_bicgstab_solve: op::bicgstab::apply (solvers/bicgstab.hh:39-187) with
 * identity (dinv == NULL) or Jacobi preconditioning, fused kernels with
 * device-resident scalars captured in HIP graphs; same operation sequence,
 * early exit (|s| < rtol, absolute, iters = iter), breakdown restart and
 * solve_info as the reference (no user diagnostic).  FP64. */
int fsx_bicgstab_solve(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv,
                       const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                       fsx_info * info);

/* fsx_gmres_solve: op::gmres::apply (solvers/gmres.hh:79-258) with identity
 * (dinv == NULL) or Jacobi preconditioning on the side s->pre_side, restart
 * length s->max_krylov_dim (<= 100; < 0 means maxiter), settings validated
 * as gmres::settings::validate (gmres.hh:394-405).  Modified Gram-Schmidt,
 * Givens rotations and the restart rule as the reference; each restart
 * cycle is one captured graph, back_solve and correct run between cycles.
 * FP64. */
int fsx_gmres_solve(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv,
                    const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                    fsx_info * info);

/* Mixed-precision iterative refinement (new, no reference counterpart):
 * outer FP64 r = b - A64 x, |r|/|b| check, inner FP32 CG on A32 to
 * `inner->rtol` with FP64 reductions, FP64 update x += |r| d.
 * jacobi: 0 inner CG unpreconditioned; 1 inner CG + Jacobi (D^{-1} of A64);
 * 2 symmetric Jacobi scaling: A32 holds S A64 S, S = D^{-1/2} (written by
 * fsx_csr_set_values_f32_sym_scaled), the inner CG runs unpreconditioned on
 * S r and d = S y -- Jacobi-preconditioned CG in exact arithmetic, without
 * the preconditioner's vector passes. */
typedef struct {
	int outer_iters;
	int inner_iters_total;
	int status;
	double rel_res_final;  /* true FP64 |b - A x| / |b| */
	double rhs_norm;
} fsx_ir_info;
int fsx_ir_cg_solve(fsx_ctx * ctx, const fsx_csr * A64, const fsx_csr * A32,
                    int jacobi, const fsx_vec * b, fsx_vec * x,
                    double rtol_outer, int max_outer,
                    const fsx_settings * inner, fsx_ir_info * info);

/* Mixed precision by residual replacement (new, no reference counterpart):
 * one FP32 CG on the symmetrically scaled system (A32 as for
 * fsx_ir_cg_solve with jacobi = 2: S A64 S in FP32 on A64's pattern) whose
 * iteration runs in segments.  A segment ends once the recursive residual
 * has fallen by `inner->rtol` (or is estimated below half the target, or
 * after inner->maxiter iterations); then x += |r0| S y in FP64, the true
 * FP64 residual r = b - A64 x is formed and checked against rtol |b|, and
 * it replaces the FP32 recursive residual, the search direction corrected
 * by the difference (p += rh_new - rh_old) and kept, so the CG continues
 * instead of restarting (residual replacement: H. A. van der Vorst and
 * Q. Ye, SIAM J. Sci. Comput. 22 (2000) 835-852).  x is the initial
 * guess.  info: outer_iters = segments, inner_iters_total = FP32
 * iterations, status converged_rtol / diverged_iters (max_segments
 * reached) / diverged_breakdown (the FP32 residual vanished while the
 * FP64 one did not), rel_res_final = true |b - A64 x| / |b|.  The same
 * pooled work vectors as fsx_ir_cg_solve, so calls on one context do not
 * overlap. */
int fsx_ir_cg_rr_solve(fsx_ctx * ctx, const fsx_csr * A64, const fsx_csr * A32,
                       const fsx_vec * b, fsx_vec * x, double rtol, int max_segments,
                       const fsx_settings * inner, fsx_ir_info * info);

/* Reference-shaped solves: the C++ op::cg / op::fcg / op::gmres /
 * op::bicgstab templates (include/flecsolve/solvers/) instantiated over the
 * device vector type and called as the reference's callers call them.
 * With the identity or Jacobi, CG, BiCGSTAB and GMRES hand the solve to the
 * fused drivers (fsx_cg_solve, fsx_cg_solve_diag when res_hist is given,
 * fsx_bicgstab_solve, fsx_gmres_solve); weighted Jacobi relaxation and FCG
 * run the generic sequence, one blocking reduction per .get() exactly as
 * the reference.  prec: 0 identity (op::I), 1 Jacobi z = D^{-1} r (Dinv
 * CSR), 2 weighted Jacobi relaxation (omega 2/3 float, `nrelax` sweeps).
 * res_hist (optional) receives the per-iteration residual estimate. */
int fsx_krylov_cg(fsx_ctx * ctx, const fsx_csr * A, int prec, int nrelax,
                  const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                  fsx_info * info, double * res_hist, int hist_cap);
/* The same CG always through the generic sequence (op::cg's generic(): the
 * reference's loop, flecsolve/solvers/cg.hh:44-140, with every vector
 * operation a libfsx kernel and three blocking reductions per iteration),
 * for comparison with the fused driver and for callers that want the
 * reference's operation order. */
int fsx_krylov_cg_generic(fsx_ctx * ctx, const fsx_csr * A, int prec, int nrelax,
                          const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                          fsx_info * info, double * res_hist, int hist_cap);
int fsx_krylov_fcg(fsx_ctx * ctx, const fsx_csr * A, int prec, int nrelax,
                   const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                   fsx_info * info);
int fsx_krylov_gmres(fsx_ctx * ctx, const fsx_csr * A, int prec, int nrelax,
                     const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                     fsx_info * info, double * res_hist, int hist_cap);
int fsx_krylov_bicgstab(fsx_ctx * ctx, const fsx_csr * A, int prec,
                        int nrelax, const fsx_vec * b, fsx_vec * x,
                        const fsx_settings * s, fsx_info * info,
                        double * res_hist, int hist_cap);

/* ---- host input path: Matrix Market and COO -> CSR ------------------------
 * matrix_market::read_header / read / parse_entry
 * (flecsolve/matrices/io/matrix_market.hh:39-103, 157-179) and coo::tocsr
 * (flecsolve/matrices/seq.hh:324-352).  Host-only; no context needed. */

/* Size line and symmetry flag; `lines` = entry lines in the file, `entries`
 * = capacity fsx_mm_read_coo needs (2 * lines for symmetric files). */
int fsx_mm_read_header(const char * path, int64_t * nrows, int64_t * ncols,
                       int64_t * lines, int * symmetric, int64_t * entries);
/* COO entries (0-based), values rounded through float as std::stof does;
 * for symmetric files each off-diagonal entry is followed by its mirror. */
int fsx_mm_read_coo(const char * path, int64_t cap, int64_t * I, int64_t * J,
                    double * V, int64_t * count);
/* Stable by row (input order inside a row), duplicates kept, columns not
 * sorted: the reference's stable_sort order, by a counting sort. */
int fsx_coo_tocsr(int64_t nrows, int64_t nnz, const int64_t * I, const int64_t * J,
                  const double * V, int64_t * rowptr, int64_t * colind, double * vals);
/* Read, convert and create the device CSR (FP64 values, index type it). */
int fsx_csr_from_mm(fsx_ctx * ctx, const char * path, fsx_itype it, fsx_csr ** out);

/* ---- adjoint of backward-Euler diffusion (new: no reference counterpart;
 * the forward step is the reference's BE integrator, time-integrators/
 * bdf.hh:74-95, 269-272, 292-296 with operator_adapter.hh:29-35, assembled
 * as A = I + (dt/h^2) L_k and solved with op::cg::apply + Jacobi). */

/* AT = A^T as a new matrix (same index and value types; rows of AT keep
 * ascending column order). */
int fsx_csr_transpose(fsx_ctx * ctx, const fsx_csr * A, fsx_csr ** AT);

/* out_vals[e] += (scale * lam[row(e)]) * x[col(e)] for every stored entry e
 * of A: the sampled outer product giving dL/dA on A's pattern. */
int fsx_csr_sampled_outer(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * lam,
                          const fsx_vec * x, double scale, fsx_vec * out_vals);

/* A's values <- shift I + scale L_k for node coefficients k on the (n+2)^3
 * lattice, 7-point stencil with arithmetic-mean faces
 * (physics/volume_diffusion/coefficient.hh:157-174); A must have the pattern
 * of fsx_gen_diffusion3d(n) (values identical to that generator's). */
int fsx_diffusion3d_values(fsx_ctx * ctx, int n, const fsx_vec * k, double shift,
                           double scale, fsx_csr * A);

/* grad_logk[p] = dL/dlog k_p given G = dL/dA on A's pattern, for A = shift I
 * + scale L_k as above. */
int fsx_diffusion3d_coeff_grad(fsx_ctx * ctx, int n, const fsx_csr * A,
                               const fsx_vec * G, const fsx_vec * k, double scale,
                               fsx_vec * grad_logk);

typedef struct {
	int forward_solves, recompute_solves, adjoint_solves;
	int failed_step; /* -1 when every solve converged */
	int64_t forward_iters, recompute_iters, adjoint_iters;
	int64_t checkpoint_bytes; /* device bytes of checkpoints + segment states */
} fsx_be_stats;

/* u^{s+1} = A^{-1} u^s for s < steps (CG + Jacobi when dinv is given, each
 * solve from the guess u^s unless s->use_zero_guess); uT = u^steps.  A solve
 * that does not converge is an error (FSX_ERR_STATE, stats->failed_step). */
int fsx_be_forward(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * dinv,
                   const fsx_vec * u0, int steps, const fsx_settings * s,
                   fsx_vec * uT, fsx_be_stats * stats);

/* Loss L = |u^steps - u_target|^2 and G = dL/dA on A's pattern (overwritten)
 * by the discrete adjoint: checkpoints every `ckpt` steps, adjoint solves
 * with AT (A itself when AT is NULL, for symmetric A) from the guess
 * lambda^{s+1}.  Optional outputs: uT = u^steps, grad_u0 = dL/du^0. */
int fsx_be_adjoint(fsx_ctx * ctx, const fsx_csr * A, const fsx_csr * AT,
                   const fsx_vec * dinv, const fsx_vec * u0,
                   const fsx_vec * u_target, int steps, int ckpt,
                   const fsx_settings * s, double * loss, fsx_vec * G,
                   fsx_vec * uT, fsx_vec * grad_u0, fsx_be_stats * stats);

/* The same with every solve (forward, recomputed and adjoint) run by
 * fsx_ir_cg_solve with jacobi = 2 from the same guess: FP32 inner CG on A32
 * (= fsx_csr_convert(A, FSX_F32) holding fsx_csr_set_values_f32_sym_scaled
 * values) to inner->rtol, FP64 outer residual and update until
 * |b - A x| / |b| <= s->rtol.  A must be symmetric (the adjoint solves use A).
 * Stats count inner iterations.  New, no reference counterpart. */
int fsx_be_forward_ir(fsx_ctx * ctx, const fsx_csr * A, const fsx_csr * A32,
                      const fsx_vec * u0, int steps, const fsx_settings * s,
                      const fsx_settings * inner, fsx_vec * uT, fsx_be_stats * stats);
int fsx_be_adjoint_ir(fsx_ctx * ctx, const fsx_csr * A, const fsx_csr * A32,
                      const fsx_vec * u0, const fsx_vec * u_target, int steps, int ckpt,
                      const fsx_settings * s, const fsx_settings * inner, double * loss,
                      fsx_vec * G, fsx_vec * uT, fsx_vec * grad_u0, fsx_be_stats * stats);
/* fsx_be_forward_rr / fsx_be_adjoint_rr: the same with every solve by
 * fsx_ir_cg_rr_solve (FP32 CG with FP64 residual replacement; inner->rtol
 * is the segment reduction, e.g. 3e-2).  New, no reference counterpart. */
int fsx_be_forward_rr(fsx_ctx * ctx, const fsx_csr * A, const fsx_csr * A32,
                      const fsx_vec * u0, int steps, const fsx_settings * s,
                      const fsx_settings * inner, fsx_vec * uT, fsx_be_stats * stats);
int fsx_be_adjoint_rr(fsx_ctx * ctx, const fsx_csr * A, const fsx_csr * A32,
                      const fsx_vec * u0, const fsx_vec * u_target, int steps, int ckpt,
                      const fsx_settings * s, const fsx_settings * inner, double * loss,
                      fsx_vec * G, fsx_vec * uT, fsx_vec * grad_u0, fsx_be_stats * stats);

/* ---- batched solves (new: many independent systems, one persistent launch)
 * A is the block-diagonal concatenation (int32 indices, FP64 values, global
 * column ids) of nsys systems; system s owns rows [sys_rows[s],
 * sys_rows[s+1]).  Each system is solved with op::cg::apply's iteration
 * (flecsolve/solvers/cg.hh:43-139) by one workgroup; per-system results are
 * written to the host arrays (any may be NULL). */
typedef struct fsx_batch fsx_batch;
int fsx_batch_create(fsx_ctx * ctx, const fsx_csr * A, int64_t nsys,
                     const int64_t * sys_rows, fsx_batch ** out);
int fsx_batch_destroy(fsx_batch * B);
/* Execution configuration: workgroup size, row mapping (0 sliced rows,
 * 1 `lanes` lanes per CSR row, 2 slice offsets). */
int fsx_batch_config(const fsx_batch * B, int * wg, int * mapping, int * lanes);
int fsx_batch_set_config(fsx_batch * B, int wg, int mapping, int lanes);
/* Autotuner: times the candidate configurations on `sample` systems for
 * `iters` CG iterations each and keeps the fastest (cached per sparsity
 * profile).  Overwrites x on the sampled systems. */
int fsx_batch_tune(fsx_ctx * ctx, fsx_batch * B, const fsx_vec * dinv,
                   const fsx_vec * b, fsx_vec * x, int sample, int iters,
                   double * best_ms);
int fsx_batched_cg(fsx_ctx * ctx, fsx_batch * B, const fsx_vec * dinv,
                   const fsx_vec * b, fsx_vec * x, const fsx_settings * s,
                   int * iters, int * status, double * res);
/* FP32-inner / FP64-outer refinement per system; A32 =
 * fsx_csr_convert(A, FSX_F32).  jacobi as for fsx_ir_cg_solve (0 none, 1
 * Jacobi-preconditioned inner CG, 2 A32 holding S A S from
 * fsx_csr_set_values_f32_sym_scaled).  outer_iters/inner_iters/status/res per
 * system. */
int fsx_batched_ir_cg(fsx_ctx * ctx, fsx_batch * B, const fsx_csr * A32,
                      int jacobi, const fsx_vec * b, fsx_vec * x,
                      double rtol_outer, int max_outer, const fsx_settings * inner,
                      int * outer_iters, int * inner_iters, int * status,
                      double * res);
/* fsx_batched_rr_cg: per system, fsx_ir_cg_rr_solve's FP32 CG with FP64
 * residual replacement (new, no reference counterpart) in the same
 * persistent launch: A32 = fsx_csr_convert(A, FSX_F32) holding
 * fsx_csr_set_values_f32_sym_scaled values; segments of inner->rtol
 * reduction (inner->maxiter iterations at most each), at most max_segments
 * per system; x is the initial guess.  The FP32 iteration is the
 * two-reduction form of the batched CG (beta = |r_new|^2 / |r|^2).
 * Outputs per system: segments, FP32 iterations, status (converged_rtol,
 * diverged_iters, diverged_breakdown) and the true FP64 |b - A x|. */
int fsx_batched_rr_cg(fsx_ctx * ctx, fsx_batch * B, const fsx_csr * A32,
                      const fsx_vec * b, fsx_vec * x, double rtol, int max_segments,
                      const fsx_settings * inner, int * segments, int * inner_iters,
                      int * status, double * res);

/* ---- measurement helpers ------------------------------------------------
 * Average duration (ms) of `reps` back-to-back launches of the SpMV kernel
 * for A, timed with HIP events on the context stream. */
int fsx_time_spmv(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * x,
                  fsx_vec * y, int reps, double * avg_ms);
/* Same for the fused CG's K1 kernel (w = A p with the per-block partials of
 * w.p, the kernel and launch shape fsx_cg_solve and the refinement's inner
 * solves use for A and p's dtype). */
int fsx_time_cg_spmv(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * p, fsx_vec * w,
                     int reps, double * avg_ms);
/* Same for the single-reduction CG's K1 (partials of w.p and w.w), which
 * the refinement's unpreconditioned inner solves (jacobi = 2) launch. */
int fsx_time_cg_spmv_sr(fsx_ctx * ctx, const fsx_csr * A, const fsx_vec * p, fsx_vec * w,
                        int reps, double * avg_ms);
/* Compulsory HBM bytes of one y = A x on the layout A's plan reads: every
 * stored value slot and index of that layout once, x read once and y written
 * once at x_bytes per element.  (For the slice-offset layout: value slots,
 * one 8-byte pointer per 64-row slice, a 4-byte offset and an 8-byte lane
 * mask per 64 slots; no column per entry.  For the symmetric stencil layout:
 * the value slots of the offsets >= 0 and one 8-byte lane mask per offset
 * and 64-row slice; the entries below the diagonal re-read stored slots and
 * are not counted again.) */
int fsx_csr_layout_bytes(const fsx_csr * A, size_t x_bytes, uint64_t * bytes);
/* Same for a plain device copy of n doubles (bandwidth reference). */
int fsx_time_copy(fsx_ctx * ctx, const fsx_vec * src, fsx_vec * dst, int reps,
                  double * avg_ms);
/* Same for fsx_parcsr_spmv (ghost copy, diag SpMV and offd rows together). */
int fsx_time_parcsr_spmv(fsx_ctx * ctx, const fsx_parcsr * M, const fsx_vec * x,
                         fsx_vec * y, int reps, double * avg_ms);

#ifdef __cplusplus
}
#endif

#endif
