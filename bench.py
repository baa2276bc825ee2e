"""Benchmark: time to relative residual 1e-10 on 256^3 variable-coefficient
diffusion, mixed precision (FP32 CG with FP64 residual replacement; the
restarted FP32-inner / FP64-outer refinement beside it) vs plain FP64 CG
(+Jacobi), with the SpMV kernel's fraction of the HBM roofline.

    python bench.py [--steps K] [--warmup W] [--n 256] [--sigma 1.0]
                    [--no-cpu-baseline]
    python bench.py --workload batched [--nsys 2048 --emin 8 --emax 48]
    python bench.py --workload adjoint [--n 128 --heat-steps 50 --ckpt 10]
    python bench.py --workload convdiff [--n 128]

A step is one complete solve from x0 = 0 to |b - A x| / |b| <= 1e-10 with the
matrix and right-hand side already resident in HBM.  The timed steps use
the mixed-precision solver (the method the headline targets; --ir-method
chooses residual replacement, the default, or the restarted refinement);
plain FP64 CG + Jacobi is timed the same way beside it.  One JSON line is
printed at the end.

`--workload batched`: 2048 independent lognormal-k diffusion systems with
edge lengths uniform in 8..48, solved in one persistent launch (FP64 CG +
Jacobi, the FP32-inner/FP64-outer refinement and the FP32 CG with FP64
residual replacement per system, autotuned row mapping); a step is the
whole batch solved to 1e-10 per system.

`--workload adjoint`: backward-Euler heat on 128^3 (dt 1e-3, 50 steps, CG +
Jacobi per step), loss |u_T - u_target|^2 and its gradient with respect to
log k by the checkpointed discrete adjoint; a step is one loss-and-gradient
evaluation; the gradient is checked against central finite differences
along 20 seeded directions.

`--workload convdiff`: 128^3 upwind convection-diffusion (cell Peclet 10)
solved with BiCGSTAB and with GMRES(30) (right Jacobi) to 1e-10.

`--workload heat`: the reference's implicit heat example (BDF2 after a CN
start, truncation-error step control, CG stage solves to 1e-6, integrated to
t = 0.1) on a 2046^2 interior grid; a step is the whole integration, run by
flecsolve_amd/heat_implicit with the fused stage path, the reference-shaped
generic path timed beside it.

`--workload parcsr [--colours 8]`: the parallel CSR product (the matrix
split over colours as the reference's parcsr splits it over ranks: ghost
copy, diag SpMV, offd rows) on the 256^3 diffusion matrix, all colours on
one device, beside the one-colour SpMV.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("time-to-1e-10 residual, 256^3 diffusion: FP32-IR vs FP64 CG; "
          "SpMV % HBM peak")
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(nrows, nnz, vbytes, ibytes=4, xbytes=8):
    """Algorithmic bytes of one y = A x: values + column ids + row pointers +
    x read once + y written once."""
    return nnz * (vbytes + ibytes) + (nrows + 1) * ibytes + nrows * xbytes * 2


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--sigma", type=float, default=1.0)
    ap.add_argument("--inner-rtol", type=float, default=1e-4,
                    help="restarted refinement: each inner solve's tolerance")
    ap.add_argument("--segment-rtol", type=float, default=3e-2,
                    help="residual replacement: reduction of the FP32 recursive "
                         "residual between two FP64 replacements")
    ap.add_argument("--ir-method", choices=["rr", "restart"], default="rr",
                    help="the timed mixed-precision solve: rr one FP32 CG with FP64 "
                         "residual replacement (fsx_ir_cg_rr_solve), restart the "
                         "restarted refinement (fsx_ir_cg_solve); the other is run once "
                         "beside it")
    ap.add_argument("--max-segments", type=int, default=100)
    ap.add_argument("--no-other-ir", action="store_true",
                    help="skip the run of the other mixed-precision method (profiling runs)")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--cpu-sample-iters", type=int, default=10,
                    help="iterations of the one-thread oracle sample")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads/colours (0: the usable CPUs, at most 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the oracle's full-size solve (CPU baseline and parity)")
    ap.add_argument("--no-sigma25", action="store_true",
                    help="skip the sigma = 2.5 comparison")
    ap.add_argument("--no-generic", action="store_true",
                    help="skip the generic-sequence CG beside the fused driver")
    ap.add_argument("--maxiter", type=int, default=5000,
                    help="CG iteration cap (FP64 solve and each inner solve)")
    ap.add_argument("--max-outer", type=int, default=10)
    ap.add_argument("--ir-jacobi", type=int, default=2, choices=[1, 2],
                    help="inner FP32 solve: 1 Jacobi-preconditioned CG on A32, "
                         "2 CG on the symmetrically Jacobi-scaled A32")
    ap.add_argument("--spmv-only", action="store_true",
                    help="time the CG K1, SpMV and copy kernels only (profiling runs)")
    ap.add_argument("--workload", choices=["diffusion", "batched", "adjoint", "convdiff",
                                           "smoother", "amg", "heat", "parcsr"],
                    default="diffusion")
    ap.add_argument("--colours", type=int, default=8,
                    help="parcsr workload: number of colours (row blocks)")
    ap.add_argument("--parcsr-spmv-only", action="store_true",
                    help="parcsr workload: time the parcsr product and the copy kernel "
                         "only (profiling runs)")
    ap.add_argument("--heat-steps", type=int, default=50)
    ap.add_argument("--ckpt", type=int, default=10)
    ap.add_argument("--fd-directions", type=int, default=20)
    ap.add_argument("--heat-rtol", type=float, default=1e-12)
    ap.add_argument("--nsys", type=int, default=2048)
    ap.add_argument("--emin", type=int, default=8)
    ap.add_argument("--emax", type=int, default=48)
    ap.add_argument("--cpu-sample-systems", type=int, default=16)
    ap.add_argument("--amg-vcycles", type=int, default=0,
                    help="amg workload: only time this many V-cycle applications "
                         "(profiling runs)")
    ap.add_argument("--amg-smoother", choices=["jacobi", "gs"], default="jacobi",
                    help="amg workload with --amg-vcycles: weighted Jacobi (2 sweeps) or "
                         "the reference's hybrid Gauss-Seidel (1 symmetric sweep)")
    ap.add_argument("--heat-n", type=int, default=2046)
    ap.add_argument("--heat-cpu-n", type=int, default=638)
    ap.add_argument("--heat-spmv-only", action="store_true",
                    help="heat workload: time the stage-matrix SpMV only (profiling runs)")
    a = ap.parse_args()
    if a.ir_method == "rr" and a.ir_jacobi != 2:
        ap.error("--ir-method rr runs on the symmetrically scaled FP32 matrix (--ir-jacobi 2)")

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the Krylov path is built and "
              "measured on a single MI355X (no sharding, no collectives); "
              "run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from flecsolve_amd import fsx
    from flecsolve_amd import problems as P

    if a.workload == "batched":
        return bench_batched(a, fsx, P)
    if a.workload == "adjoint":
        return bench_adjoint(a, fsx, P)
    if a.workload == "convdiff":
        return bench_convdiff(a, fsx, P)
    if a.workload == "smoother":
        return bench_smoother(a, fsx, P)
    if a.workload == "amg":
        return bench_amg(a, fsx, P)
    if a.workload == "heat":
        return bench_heat(a, fsx, P)
    if a.workload == "parcsr":
        return bench_parcsr(a, fsx, P)

    return bench_diffusion(a, fsx, P)


def host_cpu():
    """CPU model, logical CPUs and the CPUs this process may use."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "usable": usable}


def diffusion_problem(P, n, sigma):
    seeds = P.seeds()
    k = P.lognormal_field(n, sigma, 8.0, seeds["diffusion_k"])
    rp, ci, v = P.diffusion3d(n, k)
    b = P.uniform(n ** 3, seeds["diffusion_rhs"])
    return rp, ci, v, b


def solve_fp64(fsx, ctx, A64, dinv, bd, x, maxiter, rtol=1e-10):
    s = fsx.Settings(maxiter=maxiter, rtol=rtol, use_zero_guess=True)
    ctx.sync()
    t = time.perf_counter()
    info = fsx.cg_solve(A64, bd, x, s, dinv=dinv)
    ctx.sync()
    return info, time.perf_counter() - t


def solve_ir(fsx, ctx, A64, A32, bd, x, a, method=None):
    """The mixed-precision solve from x0 = 0: residual replacement (rr) or
    the restarted refinement (restart)."""
    method = method or a.ir_method
    fsx.check(fsx.lib().fsx_zero(ctx.h, x.h))  # x0 = 0
    ctx.sync()
    t = time.perf_counter()
    if method == "rr":
        inner = fsx.Settings(maxiter=a.maxiter, rtol=a.segment_rtol)
        ii = fsx.ir_cg_rr_solve(A64, A32, bd, x, 1e-10, inner, max_segments=a.max_segments)
    else:
        inner = fsx.Settings(maxiter=a.maxiter, rtol=a.inner_rtol)
        ii = fsx.ir_cg_solve(A64, A32, bd, x, 1e-10, inner, jacobi=a.ir_jacobi,
                             max_outer=a.max_outer)
    ctx.sync()
    return ii, time.perf_counter() - t


IR_METHODS = {
    "rr": "one FP32 CG on the symmetrically Jacobi-scaled FP32 matrix, its recursive "
          "residual replaced by the true FP64 residual (x updated in FP64) after each "
          "segment of --segment-rtol reduction, the search direction kept "
          "(fsx_ir_cg_rr_solve)",
    "restart": "restarted refinement: FP64 residual, FP32 CG correction solve to "
               "--inner-rtol from zero, FP64 update (fsx_ir_cg_solve)",
}


def true_rel_res(fsx, ctx, A64, bd, x):
    r = ctx.zeros(bd.n)
    A64.residual(bd, x, r)
    return fsx.nrm2(r) / fsx.nrm2(bd)


def rel_err(x, y):
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


def bench_diffusion(a, fsx, P):
    n = a.n
    t0 = time.perf_counter()
    rp, ci, v, b = diffusion_problem(P, n, a.sigma)
    N, nnz = n ** 3, ci.size
    log(f"problem {n}^3: {N} rows, {nnz} nnz, generated in {time.perf_counter()-t0:.1f} s")

    ctx = fsx.Context(0)
    t0 = time.perf_counter()
    A64 = ctx.csr(rp, ci, v)
    A32 = A64.convert(fsx.F32)
    if a.ir_jacobi == 2:
        A32.set_values_sym_scaled(A64)  # the refinement's inner matrix (setup)
    dinv = A64.diag_inv()
    bd = ctx.vec(b)
    x = ctx.zeros(N)
    log(f"uploaded and layouts built in {time.perf_counter()-t0:.1f} s; "
        f"SpMV plan {A64.plan()}, FP32 copy {A32.plan()}")

    # ---- kernel roofline (HIP events on the library stream) --------------
    # The timed region's dominant kernel is the inner FP32 CG's K1 on the
    # refinement's FP32 matrix: w = A p with the partials of w.p and w.w (the
    # single-reduction inner CG of the scaled matrix) or of w.p (Jacobi-
    # preconditioned inner CG); the FP64 product and the FP64 CG's K1 are
    # reported beside it.
    reps = a.spmv_reps
    p32 = ctx.vec(b.astype(np.float32))
    w32 = ctx.zeros(N, fsx.F32)
    k1_sr = a.ir_jacobi == 2
    k1_ms = fsx_time(fsx, "fsx_time_cg_spmv_sr" if k1_sr else "fsx_time_cg_spmv",
                     ctx, A32, p32, w32, reps)
    k1_bytes = A32.layout_bytes(4)
    k1_csr = spmv_bytes(N, nnz, 4, xbytes=4)
    del p32, w32
    y = ctx.zeros(N)
    spmv_ms = fsx_time(fsx, "fsx_time_spmv", ctx, A64, bd, y, reps)
    k164_ms = fsx_time(fsx, "fsx_time_cg_spmv", ctx, A64, bd, y, reps)
    b64 = A64.layout_bytes(8)
    b64_csr = spmv_bytes(N, nnz, 8)
    copy_ms = fsx_time(fsx, "fsx_time_copy", ctx, None, bd, y, reps)
    copy_gbs = 16.0 * N / (copy_ms * 1e-3) / 1e9
    del y

    def gbs(bytes_, ms):
        return bytes_ / (ms * 1e-3) / 1e9

    log(f"inner fp32 K1: {k1_ms:.4f} ms, {gbs(k1_bytes, k1_ms):.0f} GB/s on layout bytes "
        f"({gbs(k1_bytes, k1_ms) / HBM_PEAK_GBS:.1%} of 8 TB/s)")
    log(f"fp64 spmv: {spmv_ms:.4f} ms, {gbs(b64, spmv_ms):.0f} GB/s layout; fp64 K1 "
        f"{k164_ms:.4f} ms; copy kernel {copy_gbs:.0f} GB/s")
    # HBM bytes per launch measured by rocprofv3 PMC passes (FETCH_SIZE and
    # WRITE_SIZE, scaled by the copy-kernel calibration), when profiled for
    # this workload
    traffic = {}
    pmc = ROOT / "profiles" / "cg_k1_256cube_pmc.json"
    if n == 256 and abs(a.sigma - 1.0) < 1e-12 and a.ir_jacobi == 2 and pmc.exists():
        traffic = json.loads(pmc.read_text())
    knames = {0: "spmv_stream_kernel (LDS-staged CSR)", 1: "spmv_vector_kernel",
              2: "spmv_sliced_kernel (sliced rows)",
              3: "spmv_sdia_kernel (slice-offset layout)",
              4: "spmv_stencil_kernel (symmetric stencil layout)"}
    bases = {3: "compulsory bytes of the slice-offset layout (values, offsets, lane masks, "
                "slice pointers; input read once, output written once)",
             4: "compulsory bytes of the symmetric stencil layout (values at offsets >= 0 and "
                "lane masks; entries below the diagonal re-read from stored values, counted "
                "once; input read once, output written once)"}
    csr_basis = "CSR bytes (values, int32 columns, row pointers; input read once, output written once)"
    plan32, plan64 = A32.plan()[0], A64.plan()[0]
    kname, basis = knames[plan32], bases.get(plan32, csr_basis)
    roofline = {
        "bound": "hbm",
        "kernel": f"{kname} with the fused CG epilogue (K1: w = A p, partials of "
                  f"{'w.p and w.w' if k1_sr else 'w.p'}), FP32 values and vectors: the "
                  "inner CG of the refinement",
        "achieved": round(gbs(k1_bytes, k1_ms), 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(gbs(k1_bytes, k1_ms) / HBM_PEAK_GBS, 4),
        "traffic": traffic.get("k1_fp32", {}).get("traffic_bytes"),
        "algorithmic_bytes": k1_bytes,
        "bytes_basis": basis,
        "avg_ms": round(k1_ms, 5),
        # the CSR-with-int32-columns figure per product over the same time:
        # a rate for comparison with CSR kernels, not a roofline fraction
        # (this layout stores no column per entry)
        "csr_equiv_bytes": k1_csr,
        "csr_equiv_gbs": round(gbs(k1_csr, k1_ms), 1),
        "fp64_kernel": knames[plan64],
        "fp64_bytes_basis": bases.get(plan64, csr_basis),
        "fp64_spmv": {"avg_ms": round(spmv_ms, 5), "algorithmic_bytes": b64,
                      "achieved": round(gbs(b64, spmv_ms), 1),
                      "frac": round(gbs(b64, spmv_ms) / HBM_PEAK_GBS, 4),
                      "csr_equiv_bytes": b64_csr,
                      "csr_equiv_gbs": round(gbs(b64_csr, spmv_ms), 1),
                      "traffic": traffic.get("spmv_fp64", {}).get("traffic_bytes")},
        "fp64_k1": {"avg_ms": round(k164_ms, 5),
                    "achieved": round(gbs(b64, k164_ms), 1),
                    "frac": round(gbs(b64, k164_ms) / HBM_PEAK_GBS, 4),
                    "traffic": traffic.get("k1_fp64", {}).get("traffic_bytes")},
        "copy_kernel_gbs": round(copy_gbs, 1),
        "traffic_source": "profiles/cg_k1_256cube_pmc.md" if traffic else None,
    }
    # the rate of the measured HBM-side traffic over the live time: what the
    # memory system delivered (beside `achieved`, the compulsory-byte rate)
    for sub, ms in ((roofline, k1_ms), (roofline["fp64_spmv"], spmv_ms),
                    (roofline["fp64_k1"], k164_ms)):
        if sub.get("traffic"):
            sub["traffic_gbs"] = round(gbs(sub["traffic"], ms), 1)
            sub["traffic_frac"] = round(gbs(sub["traffic"], ms) / HBM_PEAK_GBS, 4)
    if a.spmv_only:
        print(json.dumps({"metric": "CG K1 / SpMV % HBM peak", "config": {
            "workload": f"diffusion3d_{n}^3", "rows": N, "nnz": int(nnz)},
            "roofline": roofline}), flush=True)
        return

    # ---- FP64 CG + Jacobi -------------------------------------------------
    cg_times = []
    for i in range(max(1, min(a.steps, 2)) + 1):
        info64, dt = solve_fp64(fsx, ctx, A64, dinv, bd, x, a.maxiter)
        cg_times.append(dt)
        log(f"fp64 cg run {i}: {dt:.3f} s, {info64.iters} iters, {info64.reason}")
        if info64.reason != "converged_rtol":
            sys.exit(f"fp64 CG did not converge within {a.maxiter} iterations")
    cg_s = min(cg_times[1:])
    x64 = x.get()
    rel64 = true_rel_res(fsx, ctx, A64, bd, x)

    # ---- FP32-inner / FP64-outer refinement (timed steps) -----------------
    for i in range(a.warmup):
        ii, dt = solve_ir(fsx, ctx, A64, A32, bd, x, a)
        log(f"ir warmup {i}: {dt:.3f} s outer {ii.outer_iters} inner {ii.inner_iters_total}")
    ctx.sync()
    times = []
    for i in range(a.steps):
        ii, dt = solve_ir(fsx, ctx, A64, A32, bd, x, a)
        times.append(dt)
        log(f"ir step {i}: {dt:.3f} s outer {ii.outer_iters} "
            f"inner {ii.inner_iters_total} rel res {ii.rel_res_final:.3e}")
        if ii.status != 1:
            sys.exit("refinement did not reach 1e-10")
    ir_s = sum(times) / len(times)
    # a caller holding b and x in host memory: one upload of b and one
    # download of x around the solve (pageable numpy arrays; the matrix is
    # uploaded once per matrix and not counted)
    ctx.sync()
    t = time.perf_counter()
    b_up = ctx.vec(b)
    ctx.sync()
    t_up = time.perf_counter() - t
    t = time.perf_counter()
    xir = x.get()
    t_down = time.perf_counter() - t
    del b_up
    rel_ir = true_rel_res(fsx, ctx, A64, bd, x)
    roofline["share_of_step"] = round(k1_ms * 1e-3 * ii.inner_iters_total / ir_s, 3)
    # the other mixed-precision method once (after a warm run), beside it
    other = "restart" if a.ir_method == "rr" else "rr"
    if a.ir_jacobi == 2 and not a.no_other_ir:
        solve_ir(fsx, ctx, A64, A32, bd, x, a, other)
        io_, dto = solve_ir(fsx, ctx, A64, A32, bd, x, a, other)
        log(f"{other}: {dto:.3f} s outer {io_.outer_iters} inner {io_.inner_iters_total} "
            f"rel res {io_.rel_res_final:.3e}")
        other_block = {"method": IR_METHODS[other], "seconds": round(dto, 4),
                       "outer_iters": io_.outer_iters,
                       "inner_iters_total": io_.inner_iters_total,
                       "true_rel_residual": true_rel_res(fsx, ctx, A64, bd, x),
                       "speedup_vs_fp64_cg": round(cg_s / dto, 3)}
    else:
        other_block = None

    # ---- the same FP64 CG + Jacobi through the generic sequence -----------
    # (the reference's loop over per-operation kernels, three blocking
    # reductions per iteration) beside the fused driver, with the host
    # synchronisations each makes
    generic = None
    if not a.no_generic:
        s0 = ctx.counters()[0]
        info_f, dt_f = solve_fp64(fsx, ctx, A64, dinv, bd, x, a.maxiter)
        syncs_f = ctx.counters()[0] - s0
        st = fsx.Settings(maxiter=a.maxiter, rtol=1e-10, use_zero_guess=True)
        fsx.check(fsx.lib().fsx_zero(ctx.h, x.h))
        ctx.sync()
        s0 = ctx.counters()[0]
        t = time.perf_counter()
        info_g, _ = fsx.krylov("cg_generic", A64, bd, x, st, prec=1)
        ctx.sync()
        dt_g = time.perf_counter() - t
        syncs_g = ctx.counters()[0] - s0
        log(f"generic cg+jacobi: {dt_g:.3f} s, {info_g.iters} iters, {syncs_g} host syncs; "
            f"fused {dt_f:.3f} s, {info_f.iters} iters, {syncs_f} host syncs")
        generic = {"seconds": round(dt_g, 4), "iters": info_g.iters,
                   "status": info_g.reason, "host_syncs": int(syncs_g),
                   "true_rel_residual": true_rel_res(fsx, ctx, A64, bd, x),
                   "fused_seconds": round(dt_f, 4), "fused_iters": info_f.iters,
                   "fused_host_syncs": int(syncs_f),
                   "fused_speedup": round(dt_g / dt_f, 3),
                   "path": "fsx_krylov_cg_generic: op::cg generic() over vec::device "
                           "and op::core<mat::device_csr> (one run, work vectors "
                           "allocated inside)"}

    out = {
        "metric": METRIC,
        "value": round(ir_s, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ir_s * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-outer",
        "data": "synthetic (seeded lognormal GRF coefficient, uniform rhs)",
        "config": {
            "workload": f"diffusion3d_{n}^3_lognormal_sigma{a.sigma:g}_rtol1e-10",
            "rows": N, "nnz": int(nnz), "index": "int32",
            "preconditioner": "jacobi",
            "ir": {"method": IR_METHODS[a.ir_method],
                   "segment_rtol" if a.ir_method == "rr" else "inner_rtol":
                       a.segment_rtol if a.ir_method == "rr" else a.inner_rtol,
                   "inner_preconditioning": ("symmetric Jacobi scaling of the FP32 matrix"
                                             if a.ir_jacobi == 2 or a.ir_method == "rr"
                                             else "Jacobi"),
                   "outer_iters": ii.outer_iters,
                   "inner_iters_total": ii.inner_iters_total,
                   "true_rel_residual": rel_ir,
                   "rel_err_vs_fp64_cg": rel_err(xir, x64)},
            "fp64_cg": {"seconds": round(cg_s, 4), "iters": info64.iters,
                        "true_rel_residual": rel64},
            "ir_speedup_vs_fp64_cg": round(cg_s / ir_s, 3),
            "host_boundary": {"upload_b_s": round(t_up, 4), "download_x_s": round(t_down, 4),
                              "pcie_inclusive_seconds": round(ir_s + t_up + t_down, 4),
                              "note": "b uploaded and x downloaded around one solve "
                                      "(pageable host arrays); value keeps them in HBM"},
            "ir_" + other: other_block,
            "generic_cg": generic,
        },
        "roofline": roofline,
    }

    # ---- a tight FP64 solve: the reference point for forward errors --------
    tight = None
    if not a.no_cpu_baseline:
        it_, dt = solve_fp64(fsx, ctx, A64, dinv, bd, x, 4 * a.maxiter, rtol=1e-13)
        tight = x.get()
        log(f"tight fp64 cg (rtol 1e-13): {it_.iters} iters, {dt:.2f} s, "
            f"true rel res {true_rel_res(fsx, ctx, A64, bd, x):.2e}")
    del A32

    # ---- the same comparison at sigma = 2.5 (one solve each) ---------------
    if not a.no_sigma25 and abs(a.sigma - 2.5) > 1e-12:
        out["config"]["sigma2_5"] = bench_sigma25(a, fsx, P, ctx, n)

    # ---- the oracle to convergence on the same problem ---------------------
    if not a.no_cpu_baseline:
        del A64, dinv, bd, x
        out["cpu_baseline"], out["config"]["oracle_parity"] = oracle_run(
            a, rp, ci, v, b, info64.iters, x64, xir, tight)
    print(json.dumps(out), flush=True)


def bench_sigma25(a, fsx, P, ctx, n):
    """FP64 CG + Jacobi and the refinement once each (after a warm run) on
    the sigma = 2.5 field: the mixed-precision configuration's contrast."""
    t0 = time.perf_counter()
    rp, ci, v, b = diffusion_problem(P, n, 2.5)
    A64 = ctx.csr(rp, ci, v)
    A32 = A64.convert(fsx.F32)
    if a.ir_jacobi == 2:
        A32.set_values_sym_scaled(A64)
    dinv = A64.diag_inv()
    bd = ctx.vec(b)
    x = ctx.zeros(n ** 3)
    del rp, ci, v
    log(f"sigma 2.5 problem set up in {time.perf_counter() - t0:.1f} s")
    cg_s, ir_s = [], []
    for i in range(2):
        info, dt = solve_fp64(fsx, ctx, A64, dinv, bd, x, 4 * a.maxiter)
        cg_s.append(dt)
        log(f"sigma 2.5 fp64 cg run {i}: {dt:.3f} s, {info.iters} iters, {info.reason}")
    x64 = x.get()
    rel64 = true_rel_res(fsx, ctx, A64, bd, x)
    for i in range(2):
        ii, dt = solve_ir(fsx, ctx, A64, A32, bd, x, a)
        ir_s.append(dt)
        log(f"sigma 2.5 ir run {i}: {dt:.3f} s outer {ii.outer_iters} "
            f"inner {ii.inner_iters_total}")
    relir = true_rel_res(fsx, ctx, A64, bd, x)
    xir = x.get()
    other = "restart" if a.ir_method == "rr" else "rr"
    other_block = None
    if a.ir_jacobi == 2:
        io_, dto = solve_ir(fsx, ctx, A64, A32, bd, x, a, other)
        log(f"sigma 2.5 {other}: {dto:.3f} s outer {io_.outer_iters} "
            f"inner {io_.inner_iters_total}")
        other_block = {"method": IR_METHODS[other], "seconds": round(dto, 4),
                       "outer_iters": io_.outer_iters,
                       "inner_iters_total": io_.inner_iters_total,
                       "true_rel_residual": true_rel_res(fsx, ctx, A64, bd, x),
                       "speedup_vs_fp64_cg": round(cg_s[-1] / dto, 3)}
    info_t, _ = solve_fp64(fsx, ctx, A64, dinv, bd, x, 8 * a.maxiter, rtol=1e-13)
    tight = x.get()
    return {
        "fp64_cg": {"seconds": round(cg_s[-1], 4), "iters": info.iters,
                    "status": info.reason, "true_rel_residual": rel64},
        "ir": {"seconds": round(ir_s[-1], 4), "outer_iters": ii.outer_iters,
               "inner_iters_total": ii.inner_iters_total,
               "status": "converged_rtol" if ii.status == 1 else "not converged",
               "true_rel_residual": relir,
               "rel_err_vs_fp64_cg": rel_err(xir, x64)},
        "ir_speedup_vs_fp64_cg": round(cg_s[-1] / ir_s[-1], 3),
        "ir_" + other: other_block,
        "forward_error_vs_tight_solve": {"fp64_cg": rel_err(x64, tight),
                                         "ir": rel_err(xir, tight),
                                         "tight_iters": info_t.iters},
    }


def oracle_run(a, rp, ci, v, b, iters_gpu, x64, xir, tight):
    """The oracle (the reference's CG + Dinv path restated on the CPU) run to
    convergence on the same problem, with P colours on P threads (the
    reference's MPI backend with one colour per rank: row blocks, block
    partials summed in block order).  Its measured time is the CPU baseline;
    its solution checks the GPU solutions at full size."""
    from oracle import oracle as O
    cpu = host_cpu()
    nthr = a.cpu_threads if a.cpu_threads > 0 else max(1, min(16, cpu["usable"]))
    Ao = O.Csr(rp, ci, v)
    dv = Ao.dinv().vals
    s = O.Settings(maxiter=a.maxiter, rtol=1e-10, use_zero_guess=True)
    log(f"oracle cg+jacobi to 1e-10: {nthr} colours on {nthr} threads "
        f"({cpu['model']}, {cpu['nproc']} CPUs, {cpu['usable']} usable) ...")
    t0 = time.perf_counter()
    xo, io = O.cg_colours(Ao, dv, b, np.zeros_like(b), s, nthr)
    dt = time.perf_counter() - t0
    log(f"oracle: {io.iters} iters ({io.reason}) in {dt:.1f} s")
    # one thread (the reference's seq path): a short sample, per iteration
    s1 = O.Settings(maxiter=a.cpu_sample_iters, rtol=1e-10, use_zero_guess=True)
    t0 = time.perf_counter()
    O.cg_colours(Ao, dv, b, np.zeros_like(b), s1, 1)
    dt1 = (time.perf_counter() - t0) / a.cpu_sample_iters
    e64, eir = rel_err(x64, xo), rel_err(xir, xo)
    fe = {"oracle": rel_err(xo, tight), "fp64_cg": rel_err(x64, tight),
          "ir": rel_err(xir, tight)}
    parity = {
        "oracle_iters": io.iters, "oracle_status": io.reason,
        "fp64_cg_iters": iters_gpu, "iters_diff": abs(iters_gpu - io.iters),
        "iters_bar": 2,
        "rel_err_fp64_cg_vs_oracle": e64, "rel_err_ir_vs_oracle": eir, "x_bar": 1e-10,
        # forward errors against an FP64 solve to rtol 1e-13: the difference
        # of two solutions is at most the sum of theirs
        "forward_error_vs_tight_solve": fe,
        "fp64_cg_within_forward_errors": e64 <= fe["oracle"] + fe["fp64_cg"],
        "ir_within_forward_errors": eir <= fe["oracle"] + fe["ir"],
    }
    base = {
        "value": round(dt, 2), "unit": "s", "cores": nthr, "kind": "port",
        "sample": (f"the whole solve: oracle CG + Jacobi FP64 on the same {len(b)}-row "
                   f"problem from x0 = 0 to rtol 1e-10, {io.iters} iterations, {nthr} "
                   f"colours on {nthr} threads"),
        "cpu_model": cpu["model"], "nproc": cpu["nproc"], "usable_cpus": cpu["usable"],
        "single_thread_ms_per_iter": round(dt1 * 1e3, 1),
        "single_thread_sample": f"{a.cpu_sample_iters} iterations on 1 thread",
    }
    return base, parity


def batch_cg_bytes(rows_s, nnz_s, vbytes, xbytes, matrix_bytes=None):
    """Algorithmic bytes of one CG + Jacobi iteration of one system: the SpMV
    (the matrix bytes, p read once, w written) and the vector passes (x, r
    read+written; p, w, dinv read; z written; then p read+written and z
    read).  matrix_bytes: per system (the layout's bytes apportioned by
    stored entries); default the CSR arrays (values + column ids + row
    pointers)."""
    if matrix_bytes is None:
        matrix_bytes = nnz_s * (vbytes + 4) + (rows_s + 1) * 4
    spmv = matrix_bytes + rows_s * xbytes * 2
    vec = rows_s * xbytes * (4 + 3 + 1 + 3)
    return spmv + vec


def bench_batched(a, fsx, P):
    seeds = P.seeds()
    t0 = time.perf_counter()
    e, rows, rp, ci, v = P.batch(a.nsys, a.emin, a.emax, seeds["batched_edges"],
                                 seeds["batched_k"], a.sigma)
    N, nnz = int(rows[-1]), int(ci.size)
    b = P.uniform(N, seeds["batched_rhs"])
    sizes = np.diff(rows)
    nnz_s = np.diff(rp[rows].astype(np.int64))
    log(f"batch {a.nsys} systems, edges {a.emin}..{a.emax}: {N} rows, {nnz} nnz, "
        f"generated in {time.perf_counter()-t0:.1f} s")
    ctx = fsx.Context(0)
    A64 = ctx.csr(rp, ci, v)
    A32 = A64.convert(fsx.F32)
    dinv = A64.diag_inv()
    B = fsx.Batch(A64, rows)
    bd, x = ctx.vec(b), ctx.zeros(N)
    log(f"uploaded; matrix plan {A64.plan()}")
    tune_ms = B.tune(bd, x, dinv, sample=64, iters=20)
    cfg = B.config()
    mapping_names = {0: "sliced", 1: "lanes", 2: "slice-offsets"}
    log(f"autotuned: workgroup {cfg[0]}, mapping {mapping_names[cfg[1]]}, "
        f"lanes {cfg[2]} ({tune_ms:.2f} ms per sample run)")
    L = fsx.lib()
    s64 = fsx.Settings(maxiter=a.maxiter, rtol=1e-10, use_zero_guess=True)
    inner = fsx.Settings(maxiter=a.maxiter, rtol=a.inner_rtol)
    r = ctx.zeros(N)

    def true_rel():
        A64.residual(bd, x, r)
        rh = r.get()
        num = np.sqrt(np.add.reduceat(rh * rh, rows[:-1]))
        den = np.sqrt(np.add.reduceat(b * b, rows[:-1]))
        return num / den

    def timed(fn, label):
        ts = []
        for i in range(a.warmup + a.steps):
            fsx.check(L.fsx_zero(ctx.h, x.h))
            ctx.sync()
            t = time.perf_counter()
            out = fn()
            ctx.sync()
            dt = time.perf_counter() - t
            log(f"{label} {'warmup' if i < a.warmup else 'step'} {i}: {dt:.3f} s")
            if i >= a.warmup:
                ts.append(dt)
        return out, sum(ts) / len(ts)

    (it64, st64, _), t64 = timed(lambda: B.cg(bd, x, s64, dinv=dinv), "fp64 batched cg")
    if not (st64 == 1).all():
        sys.exit(f"{int((st64 != 1).sum())} systems did not converge (FP64)")
    rel64 = true_rel()
    x64 = x.get()
    if a.ir_jacobi == 2:
        A32.set_values_sym_scaled(A64)  # setup, like the FP32 conversion
    (oit, iit, stir, _), tir = timed(
        lambda: B.ir_cg(A32, bd, x, 1e-10, inner, jacobi=a.ir_jacobi, max_outer=a.max_outer),
        "ir batched cg")
    if not (stir == 1).all():
        sys.exit(f"{int((stir != 1).sum())} systems did not reach 1e-10 (IR)")
    relir = true_rel()
    xir = x.get()
    err_ir = float(np.max(np.sqrt(np.add.reduceat((xir - x64) ** 2, rows[:-1])) /
                          np.sqrt(np.add.reduceat(x64 * x64, rows[:-1]))))
    # the batch matrix's layout bytes (values and index of the layout the
    # kernel reads, without the vectors) apportioned to the systems by their
    # stored entries; the CSR figure beside it
    mat_bytes = A64.layout_bytes(8) - 16 * int(rows[-1])
    mat_s = mat_bytes * nnz_s.astype(np.float64) / float(nnz_s.sum())
    bytes64 = float(np.sum(it64.astype(np.float64) * batch_cg_bytes(sizes, nnz_s, 8, 8, mat_s)))
    bytes64_csr = float(np.sum(it64.astype(np.float64) * batch_cg_bytes(sizes, nnz_s, 8, 8)))
    gbs = bytes64 / t64 / 1e9
    log(f"fp64: {t64:.3f} s, iters {it64.min()}..{it64.max()} (sum {int(it64.sum())}), "
        f"{gbs:.0f} GB/s algorithmic; ir: {tir:.3f} s, outer {oit.max()}, "
        f"inner sum {int(iit.sum())}")
    # per system, the FP32 CG with FP64 residual replacement (fsx_batched_rr_cg)
    rr = None
    if a.ir_jacobi == 2:
        seg_inner = fsx.Settings(maxiter=a.maxiter, rtol=a.segment_rtol)
        (sg, sit, sst, _), trr = timed(
            lambda: B.rr_cg(A32, bd, x, 1e-10, seg_inner, max_segments=a.max_segments),
            "rr batched cg")
        if not (sst == 1).all():
            sys.exit(f"{int((sst != 1).sum())} systems did not reach 1e-10 (RR)")
        relrr = true_rel()
        xrr = x.get()
        err_rr = float(np.max(np.sqrt(np.add.reduceat((xrr - x64) ** 2, rows[:-1])) /
                              np.sqrt(np.add.reduceat(x64 * x64, rows[:-1]))))
        log(f"rr: {trr:.3f} s, segments <= {sg.max()}, FP32 iterations sum {int(sit.sum())}")
        rr = {"seconds": round(trr, 4),
              "method": "per system: " + IR_METHODS["rr"].replace("fsx_ir_cg_rr_solve",
                                                                  "fsx_batched_rr_cg"),
              "segment_rtol": a.segment_rtol, "segments_max": int(sg.max()),
              "inner_sum": int(sit.sum()), "max_true_rel_residual": float(relrr.max()),
              "max_rel_err_vs_fp64": err_rr, "speedup_vs_fp64_cg": round(t64 / trr, 3)}
    timed_s = rr["seconds"] if (rr and a.ir_method == "rr") else tir
    out = {
        "metric": "batched CG time-to-1e-10, 2048 diffusion systems (edges 8..48)",
        "value": round(timed_s, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(timed_s * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-outer",
        "data": "synthetic (seeded per-system lognormal k, uniform rhs)",
        "config": {
            "workload": f"batched_diffusion_{a.nsys}x_edges{a.emin}-{a.emax}_rtol1e-10",
            "systems": a.nsys, "rows": N, "nnz": nnz,
            "autotuned": {"workgroup": cfg[0], "mapping": mapping_names[cfg[1]],
                          "lanes": cfg[2]},
            "fp64_cg": {"seconds": round(t64, 4), "iters_min": int(it64.min()),
                        "iters_max": int(it64.max()), "iters_sum": int(it64.sum()),
                        "max_true_rel_residual": float(rel64.max())},
            "ir": {"seconds": round(tir, 4),
                   "inner_preconditioning": ("symmetric Jacobi scaling of the FP32 matrix"
                                             if a.ir_jacobi == 2 else "Jacobi"),
                   "outer_max": int(oit.max()),
                   "inner_sum": int(iit.sum()),
                   "max_true_rel_residual": float(relir.max()),
                   "max_rel_err_vs_fp64": err_ir},
            "ir_speedup_vs_fp64_cg": round(t64 / tir, 3),
            "rr": rr,
            "value_is": "rr" if (rr and a.ir_method == "rr") else "ir",
        },
        "roofline": {"bound": "hbm", "kernel": "batched_cg_kernel<jacobi, fp64>",
                     "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes": bytes64,
                     "csr_equiv_gbs": round(bytes64_csr / t64 / 1e9, 1),
                     "note": "algorithmic bytes of every system's CG iterations (the "
                             "layout's matrix bytes apportioned by stored entries, plus "
                             "the vector passes) over the FP64 batch's wall time; small "
                             "systems run from L2"},
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batched(a, rows, rp, ci, v, b, it64, x64)
    print(json.dumps(out), flush=True)


def cpu_baseline_batched(a, rows, rp, ci, v, b, it64, x64):
    """The oracle's CG + Jacobi on a sample of the batch's systems spread over
    the size range, each solved completely; the time is scaled to the whole
    batch by the GPU's per-system rows x iterations."""
    from oracle import oracle as O
    order = np.argsort(-np.diff(rows), kind="stable")
    m = min(a.cpu_sample_systems, order.size)
    samp = order[(np.arange(m) * order.size) // m]
    t_total, work_s, worst = 0.0, 0.0, 0.0
    for s in samp:
        lo, hi = rows[s], rows[s + 1]
        e0, e1 = rp[lo], rp[hi]
        A = O.Csr((rp[lo:hi + 1] - e0).astype(np.int64), (ci[e0:e1] - lo).astype(np.int64),
                  v[e0:e1])
        t = time.perf_counter()
        xo, info, _ = O.cg(A, b[lo:hi], np.zeros(hi - lo),
                           O.Settings(maxiter=a.maxiter, rtol=1e-10, use_zero_guess=True),
                           P=O.prec(1, A.dinv()))
        t_total += time.perf_counter() - t
        work_s += float(hi - lo) * max(1, info.iters)
        worst = max(worst, float(np.linalg.norm(x64[lo:hi] - xo) / np.linalg.norm(xo)))
    work_all = float(np.sum(np.diff(rows).astype(np.float64) * np.maximum(it64, 1)))
    log(f"cpu sample: {m} systems in {t_total:.1f} s; max rel err of GPU x vs oracle {worst:.2e}")
    return {
        "value": t_total * work_all / work_s,
        "unit": "s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{m} systems spread over the size range solved completely by the "
                   f"oracle's FP64 CG + Jacobi on 1 thread ({t_total:.1f} s), scaled to "
                   f"the whole batch by rows x iterations; max rel error of the GPU's "
                   f"FP64 x vs the oracle on the sample {worst:.1e}"),
    }


def cg_iter_bytes(A):
    """Algorithmic bytes of one fused CG + Jacobi iteration on FP64 A: the
    compulsory bytes of A's product on its layout plus 88 N of vector traffic
    (update 64 N, p-update 24 N)."""
    return A.layout_bytes(8) + 88 * A.nrows


def bench_adjoint(a, fsx, P):
    from flecsolve_amd.heat import HeatProblem
    n = a.n if a.n != 256 else 128
    ctx = fsx.Context(0)
    t0 = time.perf_counter()
    prob = HeatProblem(ctx, n=n, dt=1e-3, steps=a.heat_steps, ckpt=a.ckpt,
                       rtol=a.heat_rtol)
    log(f"heat {n}^3, {a.heat_steps} steps, gamma {prob.gamma:.3f}: set up and target "
        f"trajectory in {time.perf_counter()-t0:.1f} s "
        f"({prob.target_stats.forward_iters} CG iterations)")
    theta = prob.theta0
    tf = []
    for i in range(a.warmup + a.steps):
        ctx.sync()
        t = time.perf_counter()
        L = prob.loss(theta)
        ctx.sync()
        if i >= a.warmup:
            tf.append(time.perf_counter() - t)
    tg, res = [], None
    for i in range(a.warmup + a.steps):
        ctx.sync()
        t = time.perf_counter()
        res = prob.loss_and_grad(theta)
        ctx.sync()
        dt = time.perf_counter() - t
        log(f"loss+gradient {'warmup' if i < a.warmup else 'step'} {i}: {dt:.3f} s")
        if i >= a.warmup:
            tg.append(dt)
    loss, g, gu0, st = res
    t_fwd, t_grad = sum(tf) / len(tf), sum(tg) / len(tg)
    st = st.as_dict()
    log(f"forward {t_fwd:.3f} s; loss+grad {t_grad:.3f} s; stats {st}")
    # central finite differences along seeded directions
    seeds = P.seeds()
    eps, worst, t = 1e-4, 0.0, time.perf_counter()
    for j in range(a.fd_directions):
        d = P.uniform(theta.size, seeds["heat_directions"] + j)
        fd = (prob.loss(theta + eps * d) - prob.loss(theta - eps * d)) / (2 * eps)
        err = abs(fd - g @ d) / abs(fd)
        worst = max(worst, err)
        log(f"direction {j}: fd {fd:.10e} adjoint {g @ d:.10e} rel err {err:.2e}")
    log(f"finite differences: {a.fd_directions} directions in {time.perf_counter()-t:.1f} s")
    # the same evaluation with every solve a mixed-precision refinement on the
    # symmetrically scaled FP32 copy of A (fsx_be_adjoint_ir), beside it
    prob_ir = HeatProblem(ctx, n=n, dt=1e-3, steps=a.heat_steps, ckpt=a.ckpt,
                          rtol=a.heat_rtol, refine=True)
    ti_, res_ir = [], None
    for i in range(a.warmup + a.steps):
        ctx.sync()
        t = time.perf_counter()
        res_ir = prob_ir.loss_and_grad(theta)
        ctx.sync()
        dt = time.perf_counter() - t
        log(f"refined loss+gradient {'warmup' if i < a.warmup else 'step'} {i}: {dt:.3f} s")
        if i >= a.warmup:
            ti_.append(dt)
    loss_ir, g_ir, _, st_ir = res_ir
    t_ir = sum(ti_) / len(ti_)
    refined = {"seconds": round(t_ir, 4), "speedup_vs_cg": round(t_grad / t_ir, 3),
               "loss_rel_diff": abs(loss_ir - loss) / abs(loss),
               "grad_rel_diff": float(np.linalg.norm(g_ir - g) / np.linalg.norm(g)),
               "stats": st_ir.as_dict()}
    log(f"refined solves: {t_ir:.3f} s, loss rel diff {refined['loss_rel_diff']:.2e}, "
        f"gradient rel diff {refined['grad_rel_diff']:.2e}")
    # and with every solve an FP32 CG with FP64 residual replacement
    # (fsx_be_adjoint_rr)
    del prob_ir
    prob_rr = HeatProblem(ctx, n=n, dt=1e-3, steps=a.heat_steps, ckpt=a.ckpt,
                          rtol=a.heat_rtol, refine=True, method="rr",
                          inner_rtol=a.segment_rtol)
    tr_, res_rr = [], None
    for i in range(a.warmup + a.steps):
        ctx.sync()
        t = time.perf_counter()
        res_rr = prob_rr.loss_and_grad(theta)
        ctx.sync()
        dt = time.perf_counter() - t
        log(f"residual-replacement loss+gradient {'warmup' if i < a.warmup else 'step'} {i}: "
            f"{dt:.3f} s")
        if i >= a.warmup:
            tr_.append(dt)
    loss_rr, g_rr, _, st_rr = res_rr
    t_rr = sum(tr_) / len(tr_)
    refined_rr = {"seconds": round(t_rr, 4), "speedup_vs_cg": round(t_grad / t_rr, 3),
                  "segment_rtol": a.segment_rtol,
                  "loss_rel_diff": abs(loss_rr - loss) / abs(loss),
                  "grad_rel_diff": float(np.linalg.norm(g_rr - g) / np.linalg.norm(g)),
                  "stats": st_rr.as_dict()}
    log(f"residual-replacement solves: {t_rr:.3f} s, loss rel diff "
        f"{refined_rr['loss_rel_diff']:.2e}, gradient rel diff {refined_rr['grad_rel_diff']:.2e}")
    del prob_rr
    iters = st["forward_iters"] + st["recompute_iters"] + st["adjoint_iters"]
    N, nnz = prob.N, prob.A.nnz
    gbs = iters * cg_iter_bytes(prob.A) / t_grad / 1e9
    out = {
        "metric": f"BE heat {n}^3 x {a.heat_steps} steps: loss + adjoint gradient wrt log k",
        "value": round(t_grad, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_grad * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal k fields, Gaussian-bump u0)",
        "config": {
            "workload": f"heat_be_{n}^3_dt1e-3_{a.heat_steps}steps_ckpt{a.ckpt}",
            "rows": N, "nnz": int(nnz), "solve_rtol": a.heat_rtol,
            "forward_seconds": round(t_fwd, 4),
            "adjoint_stats": st,
            "fd_check": {"directions": a.fd_directions, "eps": eps,
                         "max_rel_err": worst},
            "loss": loss,
            "refined_solves": refined,
            "residual_replacement_solves": refined_rr,
        },
        "roofline": {"bound": "hbm",
                     "kernel": f"fused CG iteration (SpMV plan {prob.A.plan()[0]} + BLAS-1)",
                     "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "note": "every CG iteration of the forward, recompute and adjoint "
                             "solves at the fused floor B_spmv + 88 N bytes (B_spmv: the "
                             "compulsory bytes of the product's layout), over the "
                             "loss+gradient wall time"},
    }
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        k = np.exp(theta)
        rp, ci, v = P.diffusion3d(n, k, scale=prob.gamma, shift=1.0)
        A = O.Csr(rp, ci, v)
        u0 = prob.u0.get()
        sample = 2
        so = O.Settings(maxiter=5000, rtol=a.heat_rtol, use_zero_guess=False)
        t = time.perf_counter()
        _, it_s = O.be_forward(A, u0, sample, so)
        dt = time.perf_counter() - t
        per_iter = dt / max(1, it_s)
        out["cpu_baseline"] = {
            "value": per_iter * iters, "unit": "s", "cores": 1, "kind": "port",
            "sample": (f"{sample} BE steps of the oracle's FP64 CG + Jacobi on 1 thread "
                       f"({it_s} iterations, {dt:.1f} s), scaled by iterations to the "
                       f"{iters} CG iterations of one loss + gradient evaluation")}
    print(json.dumps(out), flush=True)


def bench_convdiff(a, fsx, P):
    n = a.n if a.n != 256 else 128
    seeds = P.seeds()
    rp, ci, v = P.convdiff3d(n, (10.0, 5.0, 2.5))
    N, nnz = n ** 3, int(ci.size)
    b = P.uniform(N, seeds["convdiff_rhs"])
    ctx = fsx.Context(0)
    A = ctx.csr(rp, ci, v)
    bd, x = ctx.vec(b), ctx.zeros(N)
    r = ctx.zeros(N)
    results = {}
    # fused device BiCGSTAB (fsx_bicgstab_solve)
    sf = fsx.Settings(maxiter=a.maxiter, rtol=1e-10, use_zero_guess=True)
    ts = []
    for i in range(a.warmup + a.steps):
        fsx.check(fsx.lib().fsx_zero(ctx.h, x.h))
        ctx.sync()
        t = time.perf_counter()
        info = fsx.bicgstab_solve(A, bd, x, sf)
        ctx.sync()
        dt = time.perf_counter() - t
        log(f"fused bicgstab {'warmup' if i < a.warmup else 'step'} {i}: {dt:.4f} s, "
            f"{info.iters} iters, {info.reason}")
        if i >= a.warmup:
            ts.append(dt)
    A.residual(bd, x, r)
    results["bicgstab_fused"] = {"seconds": round(sum(ts) / len(ts), 4), "iters": info.iters,
                                 "restarts": info.restarts, "status": info.reason,
                                 "true_rel_residual": fsx.nrm2(r) / fsx.nrm2(bd)}
    # fused device GMRES(30) + right Jacobi (fsx_gmres_solve)
    sg = fsx.Settings(maxiter=a.maxiter, rtol=1e-10, use_zero_guess=True, max_krylov_dim=30,
                      pre_side=1, restart=True)
    dinv = A.diag_inv()
    ts = []
    for i in range(a.warmup + a.steps):
        fsx.check(fsx.lib().fsx_zero(ctx.h, x.h))
        ctx.sync()
        t = time.perf_counter()
        info = fsx.gmres_solve(A, bd, x, sg, dinv=dinv)
        ctx.sync()
        dt = time.perf_counter() - t
        log(f"fused gmres30 {'warmup' if i < a.warmup else 'step'} {i}: {dt:.4f} s, "
            f"{info.iters} iters, {info.restarts} restarts, {info.reason}")
        if i >= a.warmup:
            ts.append(dt)
    A.residual(bd, x, r)
    results["gmres30_jacobi_fused"] = {"seconds": round(sum(ts) / len(ts), 4),
                                       "iters": info.iters, "restarts": info.restarts,
                                       "status": info.reason,
                                       "true_rel_residual": fsx.nrm2(r) / fsx.nrm2(bd)}
    for name, kind, st, prec in [
            ("bicgstab_generic", "bicgstab", fsx.Settings(maxiter=a.maxiter, rtol=1e-10,
                                                  use_zero_guess=True), 0),
            ("gmres30_jacobi_generic", "gmres", fsx.Settings(maxiter=a.maxiter, rtol=1e-10,
                                                     use_zero_guess=True, max_krylov_dim=30,
                                                     pre_side=1, restart=True), 1)]:
        ts = []
        for i in range(a.warmup + a.steps):
            fsx.check(fsx.lib().fsx_zero(ctx.h, x.h))
            ctx.sync()
            t = time.perf_counter()
            info, _ = fsx.krylov(kind, A, bd, x, st, prec=prec)
            ctx.sync()
            dt = time.perf_counter() - t
            log(f"{name} {'warmup' if i < a.warmup else 'step'} {i}: {dt:.3f} s, "
                f"{info.iters} iters, {info.restarts} restarts, {info.reason}")
            if i >= a.warmup:
                ts.append(dt)
        A.residual(bd, x, r)
        relres = fsx.nrm2(r) / fsx.nrm2(bd)
        results[name] = {"seconds": round(sum(ts) / len(ts), 4), "iters": info.iters,
                         "restarts": info.restarts, "status": info.reason,
                         "true_rel_residual": relres}
    out = {
        "metric": f"convection-diffusion {n}^3 time-to-1e-10: BiCGSTAB, GMRES(30)",
        "value": results["bicgstab_fused"]["seconds"],
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(results["bicgstab_fused"]["seconds"] * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (upwind convection-diffusion, uniform rhs)",
        "config": {"workload": f"convdiff3d_{n}^3_peclet10_rtol1e-10", "rows": N,
                   "nnz": nnz, "solvers": results},
    }
    print(json.dumps(out), flush=True)


def bench_smoother(a, fsx, P):
    """Symmetric hybrid Gauss-Seidel sweeps (op::hybrid_gs, one colour) on the
    256^3 diffusion matrix; the first sweep is checked bit for bit against
    the oracle's sequential sweep on the host."""
    n = a.n
    seeds = P.seeds()
    k = P.lognormal_field(n, a.sigma, 8.0, seeds["diffusion_k"])
    rp, ci, v = P.diffusion3d(n, k)
    del k
    N, nnz = n ** 3, int(ci.size)
    b = P.uniform(N, seeds["diffusion_rhs"])
    ctx = fsx.Context(0)
    A = ctx.csr(rp, ci, v)
    t0 = time.perf_counter()
    G = fsx.HybridGS(A, ncolours=1, sweep=fsx.HybridGS.SYMMETRIC, nrelax=1)
    t_an = time.perf_counter() - t0
    lf, lb = G.levels()
    log(f"hybrid_gs {n}^3: level analysis {t_an:.2f} s, {lf} forward / {lb} backward levels")
    bd, x = ctx.vec(b), ctx.zeros(N)
    G.apply(bd, x)  # first sweep from x = 0 (captures the graph)
    ctx.sync()
    first = x.get()
    ts = []
    for i in range(a.warmup + a.steps):
        ctx.sync()
        t = time.perf_counter()
        for _ in range(10):
            G.apply(bd, x)
        ctx.sync()
        dt = (time.perf_counter() - t) / 10
        log(f"symmetric sweep {'warmup' if i < a.warmup else 'step'} {i}: {dt*1e3:.3f} ms")
        if i >= a.warmup:
            ts.append(dt)
    t_sweep = sum(ts) / len(ts)
    # the one-colour stencil on a grid runs in 8^3 tiles, one launch per tile
    # diagonal; the same sweep on the level copies (one launch per level),
    # beside it
    tiles = os.environ.get("FSX_GS_TILES", "1") != "0"
    launches = 3 * ((n + 7) // 8) - 2 if tiles else lf
    os.environ["FSX_GS_TILES"] = "0"
    try:
        G0 = fsx.HybridGS(A, ncolours=1, sweep=fsx.HybridGS.SYMMETRIC, nrelax=1)
    finally:
        if tiles:
            del os.environ["FSX_GS_TILES"]
    x0 = ctx.zeros(N)
    G0.apply(bd, x0)
    ctx.sync()
    same_paths = bool(np.array_equal(x0.get(), first))
    t = time.perf_counter()
    for _ in range(10):
        G0.apply(bd, x0)
    ctx.sync()
    t_levels = (time.perf_counter() - t) / 10
    del G0, x0
    log(f"level copies: {t_levels*1e3:.3f} ms per symmetric sweep; tiles and levels "
        f"bit-identical: {same_paths}")
    # algorithmic bytes of one direction: the SpMV's (values, column ids,
    # row pointers, x read once, x written once) plus b read once
    per_dir = spmv_bytes(N, nnz, 8) + 8 * N
    gbs = 2 * per_dir / t_sweep / 1e9
    out = {
        "metric": f"hybrid Gauss-Seidel symmetric sweep, {n}^3 diffusion",
        "value": round(t_sweep * 1e3, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_sweep * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal GRF coefficient, uniform rhs)",
        "config": {"workload": f"hybrid_gs_symmetric_{n}^3_lognormal_sigma{a.sigma:g}_1colour",
                   "rows": N, "nnz": nnz, "levels_forward": lf, "levels_backward": lb,
                   "analysis_seconds": round(t_an, 3),
                   "path": "8^3 tiles in shared memory, one launch per tile diagonal"
                           if tiles else "level copies, one launch per level",
                   "launches_per_direction": launches,
                   "us_per_launch": round(t_sweep / (2 * launches) * 1e6, 2),
                   "level_copies_ms": round(t_levels * 1e3, 3),
                   "tiles_equal_level_copies": same_paths},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes": 2 * per_dir,
                     "note": "a sweep is a chain of dependent launches (tile diagonals, or "
                             "dependency levels); latency, not bandwidth, bounds it"},
    }
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        Ao = O.Csr(rp, ci, v)
        t = time.perf_counter()
        ref = Ao.hybrid_gs(b, np.zeros(N), sweep=2, nrelax=1)
        t_cpu = time.perf_counter() - t
        same = bool(np.array_equal(ref, first))
        log(f"oracle symmetric sweep: {t_cpu:.2f} s on 1 thread; GPU bit-identical: {same}")
        out["config"]["bit_identical_to_oracle"] = same
        out["cpu_baseline"] = {"value": round(t_cpu * 1e3, 1), "unit": "ms", "cores": 1,
                               "kind": "port",
                               "sample": f"one symmetric sweep of the same {n}^3 problem "
                                         "from x = 0 (the sequential sweep of the serial "
                                         "backend)"}
    print(json.dumps(out), flush=True)


def bench_amg(a, fsx, P):
    """CG preconditioned by one UA-AMG V-cycle (mg::ua::bound_solver,
    maxiter 1) on the 256^3 diffusion problem to 1e-10, with the reference's
    hybrid Gauss-Seidel smoothing and with weighted Jacobi smoothing, beside
    the fused Jacobi-preconditioned CG."""
    import scipy.sparse as sps
    n = a.n
    seeds = P.seeds()
    k = P.lognormal_field(n, a.sigma, 8.0, seeds["diffusion_k"])
    rp, ci, v = P.diffusion3d(n, k)
    del k
    N, nnz = n ** 3, int(ci.size)
    b = P.uniform(N, seeds["diffusion_rhs"])
    S = sps.csr_matrix((v, ci, rp), shape=(N, N))
    ctx = fsx.Context(0)
    A = ctx.csr(rp, ci, v)
    bd = ctx.vec(b)
    if a.amg_vcycles > 0:
        # one V-cycle per application (the preconditioner)
        gs = a.amg_smoother == "gs"
        M = fsx.AMG(A, fsx.AmgSettings(max_levels=20, maxiter=1, atol=0.0,
                                       smoother=0 if gs else 1, nrelax=1 if gs else 2))
        x = ctx.zeros(N)
        M.apply(bd, x)
        ctx.sync()
        t = time.perf_counter()
        for _ in range(a.amg_vcycles):
            M.apply(bd, x)
        ctx.sync()
        ms = (time.perf_counter() - t) / a.amg_vcycles * 1e3
        sm = "hybrid Gauss-Seidel" if gs else "Jacobi"
        log(f"V-cycle ({sm} smoothing): {ms:.3f} ms per application, levels "
            f"{[r for r, _ in M.levels()]}")
        print(json.dumps({"metric": "UA-AMG V-cycle time", "value": round(ms, 4), "unit": "ms",
                          "config": {"workload": f"amg_vcycle_{n}^3", "rows": N,
                                     "smoother": sm, "applications": a.amg_vcycles}}),
              flush=True)
        return
    s = fsx.Settings(maxiter=a.maxiter, rtol=1e-10, use_zero_guess=True)
    results = {}
    x = ctx.zeros(N)
    ts = []
    dinv = A.diag_inv()
    for i in range(a.warmup + a.steps):
        ctx.sync()
        t = time.perf_counter()
        info = fsx.cg_solve(A, bd, x, s, dinv=dinv)
        ctx.sync()
        if i >= a.warmup:
            ts.append(time.perf_counter() - t)
    results["jacobi_cg_fused"] = {"seconds": round(sum(ts) / len(ts), 4), "iters": info.iters}
    log(f"fused Jacobi CG: {results['jacobi_cg_fused']['seconds']:.3f} s, {info.iters} iters")
    for name, smoother, nrelax in (("amg_jacobi_smoother", 1, 2), ("amg_hybrid_gs", 0, 1)):
        t0 = time.perf_counter()
        M = fsx.AMG(A, fsx.AmgSettings(max_levels=20, maxiter=1, atol=0.0, smoother=smoother,
                                       nrelax=nrelax))
        t_setup = time.perf_counter() - t0
        lv = M.levels()
        log(f"{name}: setup {t_setup:.2f} s, {len(lv)} levels, rows {[r for r, _ in lv]}")
        ts = []
        for i in range(a.warmup + a.steps):
            ctx.sync()
            t = time.perf_counter()
            info = M.cg(bd, x, s)
            ctx.sync()
            dt = time.perf_counter() - t
            log(f"{name} {'warmup' if i < a.warmup else 'step'} {i}: {dt:.3f} s, "
                f"{info.iters} iters, {info.reason}")
            if i >= a.warmup:
                ts.append(dt)
        rel = float(np.linalg.norm(b - S @ x.get()) / np.linalg.norm(b))
        results[name] = {"seconds": round(sum(ts) / len(ts), 4), "iters": info.iters,
                         "status": info.reason, "setup_seconds": round(t_setup, 2),
                         "levels": len(lv), "nrelax": nrelax, "true_rel_residual": rel}
        M.destroy()
    best = min(results["amg_jacobi_smoother"]["seconds"], results["amg_hybrid_gs"]["seconds"])
    out = {
        "metric": f"time-to-1e-10 residual, {n}^3 diffusion: UA-AMG-preconditioned CG",
        "value": results["amg_jacobi_smoother"]["seconds"],
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(results["amg_jacobi_smoother"]["seconds"] * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal GRF coefficient, uniform rhs)",
        "config": {"workload": f"amg_cg_{n}^3_lognormal_sigma{a.sigma:g}_rtol1e-10",
                   "rows": N, "nnz": nnz, "solvers": results,
                   "best_amg_speedup_vs_jacobi_cg":
                       round(results["jacobi_cg_fused"]["seconds"] / best, 3)},
    }
    print(json.dumps(out), flush=True)


def heat_runs(n, fused, repeat, cfg):
    """Run flecsolve_amd/heat_implicit; one (steps, rejects, solves, iters,
    wall seconds) per repeat."""
    import subprocess
    exe = ROOT / "flecsolve_amd" / "heat_implicit"
    if not exe.exists():
        sys.exit("flecsolve_amd/heat_implicit is not built (python -m flecsolve_amd.build)")
    cmd = [str(exe), str(cfg), str(n), "--quiet", "--repeat", str(repeat)]
    if fused:
        cmd.append("--fused")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        sys.exit("heat_implicit failed:\n" + r.stdout + r.stderr)
    runs = []
    for line in r.stdout.splitlines():
        t = line.split()
        if t and t[0] == "heat":
            kv = dict(zip(t[1::2], t[2::2]))
            runs.append((int(kv["steps"]), int(kv["rejects"]), int(kv["solves"]),
                         int(kv["iters"]), float(kv["wall"])))
    return runs


def heat_stage_spmv(a, fsx, P, n, quiet=False):
    """The stage matrix I - gamma F of the heat example on n^2 (gamma of a
    mid-run BDF2 step) and its SpMV timed with HIP events on the library
    stream; returns (ms, GB/s, algorithmic bytes, plan, nnz)."""
    N = n * n
    rp, ci, v = P.poisson2d(n)
    h = 10.0 / (n + 1)
    fv = -(1.0 / (h * h)) * v
    nnz = int(ci.size)
    ctx = fsx.Context(0)
    F = ctx.csr(rp, ci, fv)
    A = F.convert(fsx.F64)
    A.shift_scale(F, 1.0, -2e-3)
    xd = ctx.vec(P.uniform(N, 7))
    y = ctx.zeros(N)
    ms = fsx_time(fsx, "fsx_time_spmv", ctx, A, xd, y, a.spmv_reps)
    # compulsory bytes of the layout the product reads; the CSR figure (a
    # column per nonzero, which the stencil and slice layouts never read)
    # as a rate beside it
    B = A.layout_bytes(8)
    B_csr = spmv_bytes(N, nnz, 8)
    gbs = B / (ms * 1e-3) / 1e9
    plan = A.plan()
    log(f"stage spmv {n}^2: {ms:.4f} ms, {gbs:.0f} GB/s on layout bytes "
        f"({gbs / HBM_PEAK_GBS:.1%}), CSR-equivalent {B_csr / (ms * 1e-3) / 1e9:.0f} GB/s, "
        f"plan {plan}")
    if not quiet:
        # the copy kernel beside it calibrates the PMC byte counters
        copy_ms = fsx_time(fsx, "fsx_time_copy", ctx, None, xd, y, a.spmv_reps)
        log(f"copy kernel {N} doubles: {copy_ms:.4f} ms")
        print(json.dumps({"metric": "stage SpMV % HBM peak", "config": {
            "workload": f"heat2d stage matrix {n}^2", "rows": N, "nnz": nnz},
            "avg_ms": round(ms, 5), "achieved": round(gbs, 1), "algorithmic_bytes": B,
            "csr_equiv_bytes": B_csr}),
            flush=True)
    return ms, gbs, B, plan, nnz, B_csr


def bench_heat(a, fsx, P):
    """The reference's implicit heat example (examples/heat_equation/
    implicit.cc with its implicit.cfg) integrated to t = 0.1: the BDF
    integrator over device vectors with the fused stage path (csr_adapter +
    graph-captured CG), the generic path beside it, the stage SpMV against
    the HBM roofline, and the oracle (numpy integrator + C CG, one thread)
    on a smaller grid as the CPU baseline."""
    cfg = ROOT / "tests" / "golden" / "heat_implicit.cfg"
    n = a.heat_n
    N = n * n
    if a.heat_spmv_only:
        return heat_stage_spmv(a, fsx, P, n)
    reps = a.warmup + a.steps
    fused = heat_runs(n, True, reps, cfg)
    generic = heat_runs(n, False, reps, cfg)
    for name, runs in (("fused", fused), ("generic", generic)):
        for i, r in enumerate(runs):
            log(f"heat {name} {'warmup' if i < a.warmup else 'step'} {i}: {r[4]:.3f} s, "
                f"{r[0]} steps, {r[1]} rejects, {r[2]} solves, {r[3]} CG iterations")
    tf = [r[4] for r in fused[a.warmup:]]
    tg = [r[4] for r in generic[a.warmup:]]
    fs, gs = sum(tf) / len(tf), sum(tg) / len(tg)
    steps, rejects, solves, iters = fused[-1][:4]

    ms, gbs, B, plan, nnz, B_csr = heat_stage_spmv(a, fsx, P, n, quiet=True)
    traffic, traffic_src = None, None
    pmc = ROOT / "profiles" / "spmv_stage_heat_2046sq_pmc.json"
    if n == 2046 and pmc.exists():
        traffic = round(json.loads(pmc.read_text())["traffic_bytes"])
        traffic_src = "profiles/spmv_stage_heat_2046sq_pmc.md"
    out = {
        "metric": "heat example (BDF2/CN, CG stages) time to t = 0.1",
        "value": round(fs, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(fs * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (the reference example's initial state: 50 on [4,6]^2)",
        "config": {"workload": f"heat2d_{n}^2_bdf2_cn_start_t0.1", "rows": N, "nnz": nnz,
                   "time_steps": steps, "rejected": rejects, "stage_solves": solves,
                   "cg_iterations": iters,
                   "ms_per_cg_iteration": round(fs * 1e3 / iters, 4),
                   "generic_path_seconds": round(gs, 4),
                   "fused_speedup_vs_generic": round(gs / fs, 3)},
        "roofline": {"bound": "hbm", "kernel": "stage-matrix SpMV (I - gamma F), plan "
                     f"{plan[0]}", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src, "algorithmic_bytes": B,
                     "bytes_basis": "compulsory bytes of the product's layout "
                                    "(fsx_csr_layout_bytes)",
                     "csr_equiv_bytes": B_csr,
                     "csr_equiv_gbs": round(B_csr / (ms * 1e-3) / 1e9, 1),
                     "avg_ms": round(ms, 5)},
    }
    if not a.no_cpu_baseline:
        import oracle.bdf as OB
        from oracle import oracle as O
        nc = a.heat_cpu_n
        secs = OB.read_sections(cfg)
        lso = secs["linear-solver.options"]
        st = O.Settings(maxiter=int(lso["maxiter"]), rtol=float(lso["rtol"]),
                        use_zero_guess=lso["use-zero-guess"] == "true")
        rpc, cic, fvc, u0 = OB.heat2d(nc, 1.0)
        op = OB.CsrRhsOp(rpc, cic, fvc)
        solve = OB.cg_stage_solver(op, st)
        ti = OB.Integrator(OB.settings_from_section(secs["time-integrator"]), op, solve,
                           u0.size)
        t = time.perf_counter()
        OB.run(ti, u0)
        tc = time.perf_counter() - t
        log(f"cpu baseline (oracle) {nc}^2: {tc:.2f} s, {ti.get_current_step()} steps, "
            f"{solve.iters} CG iterations")
        out["cpu_baseline"] = {
            "value": round(tc, 3), "unit": "s", "cores": 1, "kind": "port",
            "sample": f"the same integration on a {nc}^2 grid ({solve.iters} CG iterations, "
                      f"{ti.get_current_step()} steps): oracle integrator (numpy) + oracle "
                      "CG (C, serial)",
            "ns_per_row_iteration": round(tc * 1e9 / (nc * nc * solve.iters), 3),
            "gpu_ns_per_row_iteration": round(fs * 1e9 / (N * iters), 4)}
    print(json.dumps(out), flush=True)


def bench_parcsr(a, fsx, P):
    """The parallel CSR product (parcsr_ops::spmv over `--colours` colours,
    all on one device: ghost copy, diag SpMV, offd rows) on the 256^3
    diffusion matrix, beside the one-colour SpMV.  The colours are FleCSI
    equal_map row blocks (z-slabs of the grid when they divide n).  Checks
    on the full problem: rows without offd entries equal the one-colour
    product bit for bit, the others agree to rounding."""
    n, ncol = a.n, a.colours
    seeds = P.seeds()
    k = P.lognormal_field(n, a.sigma, 8.0, seeds["diffusion_k"])
    rp, ci, v = P.diffusion3d(n, k)
    del k
    N, nnz = n ** 3, int(ci.size)
    b = P.uniform(N, seeds["diffusion_rhs"])
    ctx = fsx.Context(0)
    A = ctx.csr(rp, ci, v)
    t0 = time.perf_counter()
    Mp = fsx.ParCSR(A, ncol)
    t_split = time.perf_counter() - t0
    infos = [Mp.colour_info(c) for c in range(ncol)]
    ng = sum(i["ghosts"] for i in infos)
    nnz_d = sum(i["nnz_diag"] for i in infos)
    nnz_o = sum(i["nnz_offd"] for i in infos)
    log(f"parcsr {n}^3 over {ncol} colours: split in {t_split:.1f} s; {ng} ghosts, "
        f"{nnz_o} offd entries; diag plan {Mp.diag_plan()}")
    bd, y1, y2 = ctx.vec(b), ctx.zeros(N), ctx.zeros(N)
    if a.parcsr_spmv_only:
        # the parcsr product alone, then the copy kernel that calibrates the
        # PMC byte counters (profiling runs)
        ms = fsx_time(fsx, "fsx_time_parcsr_spmv", ctx, Mp, bd, y2, a.spmv_reps)
        copy_ms = fsx_time(fsx, "fsx_time_copy", ctx, None, bd, y1, a.spmv_reps)
        log(f"parcsr spmv {ms:.4f} ms; copy kernel {N} doubles {copy_ms:.4f} ms")
        print(json.dumps({"metric": "parcsr SpMV", "config": {
            "workload": f"parcsr_spmv_{n}^3_{ncol}colours"}, "avg_ms": round(ms, 5)}),
            flush=True)
        return
    A.spmv(bd, y1)
    Mp.spmv(bd, y2)
    ctx.sync()
    g1, g2 = y1.get(), y2.get()
    # rows with offd entries: those whose row has a column outside its block
    rb = np.array([c * (N // ncol) + min(c, N % ncol) for c in range(ncol + 1)])
    row_col = np.repeat(np.searchsorted(rb, np.arange(N), side="right") - 1, np.diff(rp))
    col_col = np.searchsorted(rb, ci, side="right") - 1
    boundary = np.zeros(N, bool)
    boundary[np.repeat(np.arange(N), np.diff(rp))[row_col != col_col]] = True
    interior_exact = bool(np.array_equal(g1[~boundary], g2[~boundary]))
    rel = float(np.max(np.abs(g1[boundary] - g2[boundary]) /
                       np.maximum(np.abs(g1[boundary]), 1e-300))) if boundary.any() else 0.0
    log(f"check: interior rows bit-identical {interior_exact}; boundary rows "
        f"({int(boundary.sum())}) max rel diff {rel:.2e}")
    del row_col, col_col
    ms_plain = fsx_time(fsx, "fsx_time_spmv", ctx, A, bd, y1, a.spmv_reps)
    ts = []
    for i in range(a.warmup + a.steps):
        ms = fsx_time(fsx, "fsx_time_parcsr_spmv", ctx, Mp, bd, y2, a.spmv_reps)
        log(f"parcsr spmv {'warmup' if i < a.warmup else 'step'} {i}: {ms:.4f} ms")
        if i >= a.warmup:
            ts.append(ms)
    ms = sum(ts) / len(ts)
    # bytes the three launches must move: the diag blocks' SpMV (int32
    # indices), the ghost gather (colmap, x, ghost), the offd rows (row ids,
    # row pointers, columns, values, ghost values, y read and written)
    B_diag = Mp.diag_layout_bytes(8)  # compulsory bytes of its layout
    B_diag_csr = spmv_bytes(N, nnz_d, 8)
    B_gather = ng * (4 + 8 + 8)
    nb = int(boundary.sum())
    B_offd = (2 * nb + 1) * 4 + nnz_o * (4 + 8 + 8) + nb * 16
    B = B_diag + B_gather + B_offd
    gbs = B / (ms * 1e-3) / 1e9
    B_plain = A.layout_bytes(8)  # the one-colour product on its layout
    traffic, traffic_src = None, None
    pmc = ROOT / "profiles" / "parcsr_spmv_256cube_8colours_pmc.json"
    if n == 256 and ncol == 8 and abs(a.sigma - 1.0) < 1e-12 and pmc.exists():
        traffic = round(json.loads(pmc.read_text())["traffic_bytes"])
        traffic_src = "profiles/parcsr_spmv_256cube_8colours_pmc.md"
    out = {
        "metric": f"parallel CSR SpMV over {ncol} colours on one device, {n}^3 diffusion",
        "value": round(ms, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal GRF coefficient, uniform x)",
        "config": {"workload": f"parcsr_spmv_{n}^3_{ncol}colours", "rows": N, "nnz": nnz,
                   "colours": ncol, "ghosts": ng, "offd_entries": nnz_o, "offd_rows": nb,
                   "split_seconds": round(t_split, 2),
                   "one_colour_spmv_ms": round(ms_plain, 4),
                   "one_colour_gbs": round(B_plain / (ms_plain * 1e-3) / 1e9, 1),
                   "overhead_vs_one_colour": round(ms / ms_plain - 1, 4),
                   "interior_rows_bit_identical": interior_exact,
                   "boundary_rows_max_rel_diff": rel},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes": B,
                     "bytes": {"diag_spmv": B_diag, "ghost_gather": B_gather,
                               "offd_rows": B_offd},
                     "bytes_basis": "diag blocks: compulsory bytes of their layout "
                                    "(fsx_csr_layout_bytes); ghost gather and offd rows: "
                                    "CSR arrays they read and write",
                     "csr_equiv_gbs": round((B_diag_csr + B_gather + B_offd) /
                                            (ms * 1e-3) / 1e9, 1),
                     "avg_ms": round(ms, 5)},
    }
    if not a.no_cpu_baseline:
        # the oracle's colour-by-colour product on the 128^3 problem of the
        # same kind, scaled to this one by stored entries
        from oracle import parcsr as PC
        m = min(n, 128)
        km = P.lognormal_field(m, a.sigma, 8.0, seeds["diffusion_k"])
        rps, cis, vs = P.diffusion3d(m, km)
        blocks = PC.split(rps, cis, vs, m ** 3, m ** 3, ncol)
        xs = P.uniform(m ** 3, 3)
        reps, t = 3, time.perf_counter()
        for _ in range(reps):
            PC.spmv(blocks, xs, m ** 3)
        t_cpu = (time.perf_counter() - t) / reps * (nnz / cis.size)
        out["cpu_baseline"] = {"value": round(t_cpu * 1e3, 1), "unit": "ms", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle parcsr product on {m}^3 over {ncol} colours "
                                         f"(numpy, 1 thread), {reps} runs, scaled by stored "
                                         "entries"}
        log(f"oracle parcsr product: {t_cpu*1e3:.1f} ms scaled to {n}^3")
    print(json.dumps(out), flush=True)


def fsx_time(fsx, name, ctx, A, x, y, reps):
    import ctypes as C
    ms = C.c_double()
    if A is not None:
        fsx.check(getattr(fsx.lib(), name)(ctx.h, A.h, x.h, y.h, reps, C.byref(ms)), name)
    else:
        fsx.check(getattr(fsx.lib(), name)(ctx.h, x.h, y.h, reps, C.byref(ms)), name)
    return ms.value


if __name__ == "__main__":
    main()
