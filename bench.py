#!/usr/bin/env python3
"""Benchmark of the blocked SPD solver path on one MI355X.

Step (timed K times after W warm-ups) = the reference's Cholesky solve path
on the N = 65536 GP system, inputs resident in HBM:
  RBF assembly of K (N x N, packed lower block triangle, b = 256) from X
  -> FP64 blocked Cholesky K = L L^T -> forward + backward triangular solves.
value = (N^3/3) / step time in TFLOP/s (whole step, assembly and solves
included). After the timed region, untimed checks and side measurements:
backward error of the solution, per-phase kernel times of the factorisation
(HIP events, for the roofline of the dominant kernel), the mixed-precision
factor with FP64 refinement at the same backward-error target, the CG
workload (N = 32768, b = 128) with its mat-vec bandwidth, the GP likelihood
gradient (N = 16384, b = 128) checked against finite differences, and the CPU
baseline (the oracle's restatement of the reference's CPU path on a bounded
sample).

Prints one JSON line.
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "heterogeneous-solvers_amd")

METRIC = "Cholesky N=65536 solve time+TFLOP/s: FP64 vs mixed+IR at equal residual; CG GB/s"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the dominant kernel (trailing_mfma_f64<128, 0>,
# average over the look-ahead rest-update launches at N = 65536, b = 256):
# rocprofv3 FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, the factors that make the
# counters read 1 GiB for a copy kernel of 1 GiB read + 1 GiB written
# (bench.py --copy-calibration); four-panel steps, 62 launches. Summary and
# commands: profiles/trailing_mfma_f64_cholesky65536_hbm_traffic.md
TRAILING_TRAFFIC_BYTES = 23.58e9


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def packed_len(N, b):
    n = -(-N // b)
    return n * (n + 1) // 2 * b * b


def rest_update_work(hs, N, b):
    cnt, fl = C.c_int64(), C.c_double()
    hs.call("hs_cholesky_schedule_f64", N, b, C.byref(cnt), C.byref(fl))
    return cnt.value, fl.value


def mixed_update_work(hs, N, b, variant):
    cnt, fl = C.c_int64(), C.c_double()
    hs.call("hs_cholesky_mixed_schedule", N, b, variant, C.byref(cnt), C.byref(fl))
    return cnt.value, fl.value


def cholesky_workload(hs, dev, syn, args):
    seed, N, d, b, ell, sf2, sn2 = syn.WORKLOADS["cholesky"]
    if args.N:
        N = args.N
    n = -(-N // b)
    X, y = syn.gp_inputs(seed, N, d)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    dX = dev.upload(X.reshape(-1))
    dy = dev.upload(rhs)
    dx = dev.zeros(n * b)
    plen = packed_len(N, b)
    log("cholesky workload N=%d b=%d d=%d: packed K %.2f GB" % (N, b, d, plen * 8 / 1e9))
    dA = dev.alloc(plen * 8)
    t_all, t_asm, t_fac, t_sol = hs.Timer(dev), hs.Timer(dev), hs.Timer(dev), hs.Timer(dev)
    phase = (C.c_double * 3)()
    launches = (C.c_int64 * 3)()
    step_ms, phases, parts = [], [], []
    for s in range(args.warmup + args.steps):
        t_all.start()
        t_asm.start()
        hs.call("hs_rbf_assemble_f64", dev.h, dX.p, N, d, b, ell, sf2, sn2, dA.p)
        t_asm.stop()
        t_fac.start()
        info = hs.call("hs_cholesky_f64_timed", dev.h, dA.p, N, b, phase, launches)
        t_fac.stop()
        t_sol.start()
        hs.call("hs_memcpy_dtod", dev.h, dx.p, dy.p, n * b * 8)
        hs.call("hs_tri_solve_f64", dev.h, dA.p, dx.p, N, b)
        t_sol.stop()
        t_all.stop()
        ms = t_all.ms()
        part = (t_asm.ms(), t_fac.ms(), t_sol.ms())
        if info != 0:
            raise RuntimeError("factorisation reported non-positive pivot %d" % info)
        if s >= args.warmup:
            step_ms.append(ms)
            phases.append(list(phase))
            parts.append(part)
        log("  step %d: %.1f ms (assembly %.1f factor %.1f solves %.1f; kernel sums potrf %.1f trsm %.1f "
            "update %.1f ms)" % (s, ms, part[0], part[1], part[2], phase[0], phase[1], phase[2]))
    x = dev.download(dx, n * b)
    # untimed check: residual on a freshly assembled K
    hs.call("hs_rbf_assemble_f64", dev.h, dX.p, N, d, b, ell, sf2, sn2, dA.p)
    dr = dev.zeros(n * b)
    hs.call("hs_symv_packed_f64", dev.h, dA.p, dx.p, dr.p, N, b)
    Kx = dev.download(dr, n * b)
    ones = np.zeros(n * b)
    ones[:N] = 1.0
    d1 = dev.upload(ones)
    hs.call("hs_symv_packed_f64", dev.h, dA.p, d1.p, dr.p, N, b)
    knorm_inf = float(np.max(dev.download(dr, n * b)))  # RBF entries are >= 0
    res = rhs - Kx
    rel_res = float(np.linalg.norm(res) / np.linalg.norm(rhs))
    bwd = float(np.max(np.abs(res)) / (knorm_inf * np.max(np.abs(x)) + np.max(np.abs(rhs))))
    # untimed: the same system through the reference-order factor and solves
    # (the parity mode, FMA contracted), against the fast path's solution
    ref_order = None
    if not args.no_ref_order:
        dR = dev.alloc(plen * 8)
        hs.call("hs_rbf_assemble_f64", dev.h, dX.p, N, d, b, ell, sf2, sn2, dR.p)
        dxr = dev.upload(rhs)
        t_ref = hs.Timer(dev)
        t_ref.start()
        info_r = hs.call("hs_cholesky_ref_order_f64", dev.h, dR.p, N, b, 1)
        hs.call("hs_tri_solve_ref_order_f64", dev.h, dR.p, dxr.p, N, b, 1)
        t_ref.stop()
        ref_ms = t_ref.ms()
        dR.free()
        xr = dev.download(dxr, n * b)
        hs.call("hs_symv_packed_f64", dev.h, dA.p, dxr.p, dr.p, N, b)
        rr = rhs - dev.download(dr, n * b)
        ref_order = {"ms": ref_ms, "info": info_r, "fma": 1,
                     "rel_residual": float(np.linalg.norm(rr) / np.linalg.norm(rhs)),
                     "rel_diff_vs_fast": float(np.linalg.norm(xr - x) / np.linalg.norm(xr))}
        log("  reference-order factor + solves: %.0f ms, relative residual %.3e, fast path differs by %.3e"
            % (ref_ms, ref_order["rel_residual"], ref_order["rel_diff_vs_fast"]))
    dA.free()
    ph = np.array(phases).mean(axis=0)
    pt = np.array(parts).mean(axis=0)
    flops = N ** 3 / 3.0
    ms = float(np.median(step_ms))
    nl = int(launches[2])
    # algorithmic flops of the timed trailing-update launches (the look-ahead
    # rest updates): b^3 per diagonal and 2 b^3 per off-diagonal block and panel
    sl, upd_flops = rest_update_work(hs, N, b)
    if sl != nl:
        raise RuntimeError("schedule reports %d rest-update launches, timer saw %d" % (sl, nl))
    achieved = upd_flops / (ph[2] / 1e3) / 1e12
    return {
        "N": N, "b": b, "d": d, "seed": seed, "ell": ell, "sf2": sf2, "sn2": sn2,
        "ms_per_step": ms, "step_ms": step_ms, "tflops_step": flops / (ms / 1e3) / 1e12,
        "kernel_sum_ms": {"potrf": float(ph[0]), "trsm": float(ph[1]), "update": float(ph[2])},
        "assembly_ms": float(pt[0]), "factor_ms": float(pt[1]), "solve_ms": float(pt[2]),
        "tflops_factor": flops / (pt[1] / 1e3) / 1e12,
        "update": {"launches": nl, "flops_per_launch": upd_flops / nl, "avg_launch_ms": float(ph[2]) / nl,
                   "tflops": achieved},
        "rel_residual": rel_res, "backward_error_inf": bwd,
        "reference_order": ref_order,
        "_x": x[:N],  # the FP64 solution, for the mixed variants' forward difference (not printed)
    }


def mixed_workload(hs, dev, syn, args, x64=None, rr64=None):
    """Mixed precision on the same N = 65536 system: time to solution of an
    FP32-stored factor (FP32-MFMA or BF16 hi+lo trailing updates) plus FP64
    refinement, K assembled beforehand, to two targets: the normwise backward
    error 1e-12, and the FP64 path's own relative residual rr64 (equal
    residual) together with that backward error. A refinement that misses its
    target fails the run. With the FP64 path's solution x64, each variant's
    relative forward difference from it is reported."""
    seed, N, d, b, ell, sf2, sn2 = syn.WORKLOADS["cholesky"]
    if args.N:
        N = args.N
    n = -(-N // b)
    X, y = syn.gp_inputs(seed, N, d)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    plen = packed_len(N, b)
    dX, dy, dx = dev.upload(X.reshape(-1)), dev.upload(rhs), dev.zeros(n * b)
    dK = dev.alloc(plen * 8)
    dL = dev.alloc(plen * 4)
    hs.call("hs_rbf_assemble_f64", dev.h, dX.p, N, d, b, ell, sf2, sn2, dK.p)
    t_fac, t_ref = hs.Timer(dev), hs.Timer(dev)
    phase = (C.c_double * 3)()
    launches = (C.c_int64 * 3)()
    targets = [("backward_error_1e-12", 1e-12, 0.0)]
    if rr64:
        targets.append(("equal_residual", 1e-12, rr64))
    out = {}
    for name, variant in (("fp32", hs.HS_MIXED_FP32), ("bf16x3", hs.HS_MIXED_BF16X3),
                          ("f16x3", hs.HS_MIXED_F16X3)):
        runs = []
        for s in range(args.warmup + args.steps):
            t_fac.start()
            info = hs.call("hs_cholesky_mixed_timed", dev.h, dK.p, dL.p, N, b, variant, 0.0, phase, launches)
            t_fac.stop()
            if info != 0:
                raise RuntimeError("mixed factorisation reported non-positive pivot %d" % info)
            fac = t_fac.ms()
            per_target = {}
            for tname, tol, rrtol in targets:
                rep = hs.RefineReport()
                t_ref.start()
                hs.call("hs_refine_solve_to", dev.h, dK.p, dL.p, dy.p, dx.p, N, b, tol, rrtol, 60, 600,
                        C.byref(rep))
                t_ref.stop()
                per_target[tname] = (t_ref.ms(), rep.as_dict(), dev.download(dx, n * b)[:N] if x64 is not None
                                     else None)
            log("  mixed %s step %d: factor %.1f ms; refine %s" % (
                name, s, fac, {k: "%.1f ms (%s)" % (v[0], v[1]) for k, v in per_target.items()}))
            if s >= args.warmup:
                runs.append((fac, list(phase), per_target))
        fac = float(np.median([r[0] for r in runs]))
        ph = np.mean([r[1] for r in runs], axis=0)
        _, upd_flops = mixed_update_work(hs, N, b, variant)
        rec = {"factor_ms": fac, "update_kernel_ms": float(ph[2]),
               "update_tflops": upd_flops / (ph[2] / 1e3) / 1e12,
               "kernel_sum_ms": {"potrf": float(ph[0]), "trsm": float(ph[1]), "update": float(ph[2])}}
        for tname, tol, rrtol in targets:
            ref = float(np.median([r[2][tname][0] for r in runs]))
            last = runs[-1][2][tname]
            fwd = None
            if x64 is not None:
                fwd = float(np.linalg.norm(last[2] - x64) / np.linalg.norm(x64))
            rec[tname] = {"target_backward_error": tol, "target_relative_residual": rrtol or None,
                          "refine_ms": ref, "solve_ms": fac + ref, "report": last[1], "forward_diff_vs_fp64": fwd}
        out[name] = rec
    dK.free()
    dL.free()
    return out


def cg_workload(hs, dev, syn, args, symv_only=False):
    seed, N, d, b, ell, sf2, sn2 = syn.WORKLOADS["cg"]
    n = -(-N // b)
    X, y = syn.gp_inputs(seed, N, d)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    plen = packed_len(N, b)
    dX, drhs, dx = dev.upload(X.reshape(-1)), dev.upload(rhs), dev.zeros(n * b)
    dA = dev.alloc(plen * 8)
    hs.call("hs_rbf_assemble_f64", dev.h, dX.p, N, d, b, ell, sf2, sn2, dA.p)
    timer = hs.Timer(dev)
    # mat-vec alone
    dq = dev.zeros(n * b)
    hs.call("hs_symv_packed_f64", dev.h, dA.p, dx.p, dq.p, N, b)
    reps = 20
    timer.start()
    for _ in range(reps):
        hs.call("hs_symv_packed_f64", dev.h, dA.p, drhs.p, dq.p, N, b)
    timer.stop()
    symv_ms = timer.ms() / reps
    passes = (C.c_double * 2)()
    hs.call("hs_symv_packed_f64_timed", dev.h, dA.p, drhs.p, dq.p, N, b, reps, passes)
    bytes_symv = plen * 8 + 16 * n * b
    # the same packed bytes read once by a plain streaming-read kernel: the
    # read rate a bandwidth-bound pass over this buffer can reach
    dpart = dev.zeros(-(-plen // 16384))
    hs.call("hs_read_sum_f64", dev.h, dA.p, plen, dpart.p)
    timer.start()
    for _ in range(reps):
        hs.call("hs_read_sum_f64", dev.h, dA.p, plen, dpart.p)
    timer.stop()
    read_ms = timer.ms() / reps
    read_gbs = plen * 8 / (read_ms / 1e3) / 1e9
    if symv_only:
        dA.free()
        return {"N": N, "b": b, "symv_ms": symv_ms, "symv_gbs": bytes_symv / (symv_ms / 1e3) / 1e9,
                "symv_block_pass_ms": passes[0], "symv_reduce_pass_ms": passes[1],
                "symv_block_pass_gbs": bytes_symv / (passes[0] / 1e3) / 1e9,
                "stream_read_ms": read_ms, "stream_read_gbs": read_gbs}
    # the reference-order mat-vec of the parity mode: each row summed in the
    # reference's order, so the off-diagonal blocks are read twice (once
    # row-wise, once transposed): 2 B_A - the diagonal blocks' bytes
    hs.call("hs_symv_ref_order_f64", dev.h, dA.p, drhs.p, dq.p, N, b, 1)
    timer.start()
    for _ in range(5):
        hs.call("hs_symv_ref_order_f64", dev.h, dA.p, drhs.p, dq.p, N, b, 1)
    timer.stop()
    ref_symv_ms = timer.ms() / 5
    ref_symv_bytes = 2 * plen * 8 - n * b * b * 8
    it, df, d0 = C.c_int64(), C.c_double(), C.c_double()
    timer.start()
    hs.call("hs_cg_f64", dev.h, dA.p, drhs.p, dx.p, N, b, 1e-10, 100000, C.byref(it), C.byref(df), C.byref(d0))
    timer.stop()
    cg_ms = timer.ms()
    iters = it.value
    nsymv = iters + -(-iters // 50)
    x = dev.download(dx, n * b)
    hs.call("hs_symv_packed_f64", dev.h, dA.p, dx.p, dq.p, N, b)
    rel_res = float(np.linalg.norm(rhs - dev.download(dq, n * b)) / np.linalg.norm(rhs))
    dA.free()
    log("cg workload N=%d b=%d: %d iterations in %.1f ms, symv %.3f ms (%.0f GB/s); block pass %.3f ms "
        "(%.0f GB/s), reduction pass %.3f ms; plain streaming read of the same bytes %.3f ms (%.0f GB/s)"
        % (N, b, iters, cg_ms, symv_ms, bytes_symv / symv_ms / 1e6, passes[0], bytes_symv / passes[0] / 1e6,
           passes[1], read_ms, read_gbs))
    return {
        "N": N, "b": b, "seed": seed, "eps": 1e-10, "iterations": iters, "ms": cg_ms,
        "symv_count": nsymv, "symv_ms": symv_ms, "symv_gbs": bytes_symv / (symv_ms / 1e3) / 1e9,
        "symv_block_pass_ms": passes[0], "symv_reduce_pass_ms": passes[1],
        "symv_block_pass_gbs": bytes_symv / (passes[0] / 1e3) / 1e9,
        "symv_block_pass_hbm_frac": bytes_symv / (passes[0] / 1e3) / 1e9 / HBM_PEAK_GBS,
        "stream_read_ms": read_ms, "stream_read_gbs": read_gbs,
        "symv_block_pass_vs_stream_read": read_ms / passes[0],
        "cg_gbs": nsymv * bytes_symv / (cg_ms / 1e3) / 1e9, "rel_residual_true": rel_res,
        "ref_order_symv_ms": ref_symv_ms, "ref_order_symv_gbs": ref_symv_bytes / (ref_symv_ms / 1e3) / 1e9,
        "delta_ratio": df.value / d0.value if d0.value else 0.0,
    }


def adjoint_workload(hs, dev, syn, args):
    """GP NLL and its gradient (log ell, log sf2, log sn2) at N = 16384, d = 8,
    b = 128, ell = 1, sf2 = 1, sn2 = 1e-2: RBF assembly, FP64 factor, both
    solves, the reverse pass through the blocked Cholesky and the contraction
    with dK/dtheta. Checked against central differences (h = 1e-5)."""
    import math
    seed, N, d, b, ell, sf2, sn2 = syn.WORKLOADS["adjoint"]
    n = -(-N // b)
    X, y = syn.gp_inputs(seed, N, d)
    dX, dy = dev.upload(X.reshape(-1)), dev.upload(y)
    plen = packed_len(N, b)
    dK, dKb = dev.alloc(plen * 8), dev.alloc(plen * 8)
    nll, g = C.c_double(), (C.c_double * 3)()

    def grad_call():
        hs.call("hs_gp_nll_grad_f64", dev.h, dX.p, N, d, b, ell, sf2, sn2, dy.p, dK.p, dKb.p, C.byref(nll), g)

    def nll_at(th):
        v = C.c_double()
        hs.call("hs_gp_nll_f64", dev.h, dX.p, N, d, b, math.exp(th[0]), math.exp(th[1]), math.exp(th[2]), dy.p,
                dK.p, C.byref(v))
        return v.value

    grad_call()
    times = []
    for _ in range(max(1, args.steps)):
        t0 = time.perf_counter()
        grad_call()
        times.append((time.perf_counter() - t0) * 1e3)
    t0 = time.perf_counter()
    th = [math.log(ell), math.log(sf2), math.log(sn2)]
    nll_at(th)
    fwd_ms = (time.perf_counter() - t0) * 1e3
    gv = [g[0], g[1], g[2]]
    h = 1e-5
    fd = []
    for i in range(3):
        tp, tm = list(th), list(th)
        tp[i] += h
        tm[i] -= h
        fd.append((nll_at(tp) - nll_at(tm)) / (2 * h))
    rel = max(abs(a - c) for a, c in zip(gv, fd)) / max(abs(a) for a in gv)
    bwd_flops = sum(2.0 * m * m * b ** 3 + 4.0 * m * b ** 3 + 6.0 * b ** 3 for m in range(n))
    ms = float(np.median(times))
    dK.free()
    dKb.free()
    log("adjoint workload N=%d b=%d d=%d: nll+grad %.1f ms (nll alone %.1f ms), grad %s, fd %s, rel diff %.2e" %
        (N, b, d, ms, fwd_ms, gv, fd, rel))
    return {"N": N, "b": b, "d": d, "seed": seed, "ell": ell, "sf2": sf2, "sn2": sn2, "nll": nll.value,
            "grad": gv, "fd_grad": fd, "fd_h": h, "fd_max_rel_diff": rel, "ms": ms, "nll_only_ms": fwd_ms,
            "flops": N ** 3 / 3.0 + bwd_flops, "tflops": (N ** 3 / 3.0 + bwd_flops) / (ms / 1e3) / 1e12}


def gpr_workload(hs, dev, syn, args):
    """GP prediction on the CG workload's training inputs (N = 32768, d = 8,
    ell = 1, sf2 = 1) at M = 8192 test inputs (seed 102): the predictive mean
    K* alpha, fused (hs_gp_predict_f64: K* is never stored) and as the
    reference's two steps (hs_rbf_cross_f64 writes K*, hs_gemv_rows_f64 reads
    it). alpha is the training target vector: the cost does not depend on it."""
    seed, N, d, b, ell, sf2, sn2 = syn.WORKLOADS["cg"]
    M = 8192
    X, y = syn.gp_inputs(seed, N, d)
    Xt, _ = syn.gp_inputs(seed + 100, M, d)
    dX, dXt, da = dev.upload(X.reshape(-1)), dev.upload(Xt.reshape(-1)), dev.upload(y)
    dm, dm2, dK = dev.zeros(M), dev.zeros(M), dev.alloc(N * M * 8)
    timer = hs.Timer(dev)
    reps = 5

    def fused():
        hs.call("hs_gp_predict_f64", dev.h, dX.p, N, dXt.p, M, d, ell, sf2, da.p, dm.p)

    def assemble():
        hs.call("hs_rbf_cross_f64", dev.h, dX.p, N, dXt.p, M, d, ell, sf2, dK.p)

    def gemv():
        hs.call("hs_gemv_rows_f64", dev.h, dK.p, da.p, dm2.p, N, M)

    out = {"N": N, "M": M, "d": d, "ell": ell, "sf2": sf2}
    for name, f in (("fused", fused), ("kstar_assembly", assemble), ("kstar_gemv", gemv)):
        f()
        timer.start()
        for _ in range(reps):
            f()
        timer.stop()
        out[name + "_ms"] = timer.ms() / reps
    timer.close()
    a, c = dev.download(dm, M), dev.download(dm2, M)
    dK.free()
    pairs = float(N) * M
    kbytes = pairs * 8
    out["pairs_per_s_fused"] = pairs / (out["fused_ms"] / 1e3)
    out["two_step_ms"] = out["kstar_assembly_ms"] + out["kstar_gemv_ms"]
    out["kstar_assembly_gbs"] = kbytes / (out["kstar_assembly_ms"] / 1e3) / 1e9
    out["kstar_gemv_gbs"] = kbytes / (out["kstar_gemv_ms"] / 1e3) / 1e9
    out["fused_vs_two_step_max_rel_diff"] = float(np.max(np.abs(a - c)) / np.max(np.abs(c)))
    log("gpr workload N=%d M=%d d=%d: fused %.2f ms (%.3g pairs/s); two steps %.2f ms (K* %.2f ms = %.0f GB/s, "
        "mat-vec %.2f ms = %.0f GB/s); max rel diff %.1e" %
        (N, M, d, out["fused_ms"], out["pairs_per_s_fused"], out["two_step_ms"], out["kstar_assembly_ms"],
         out["kstar_assembly_gbs"], out["kstar_gemv_ms"], out["kstar_gemv_gbs"], out["fused_vs_two_step_max_rel_diff"]))
    return out


def cpu_baseline(args):
    """The oracle (C/OpenMP restatement of the reference's share-0 CPU path)
    on the host cores, one entry per workload:
      * cholesky_4096: the CPU-reference-path workload exactly (N = 4096,
        b = 128, d = 8, l = 1, sn2 = 1e-2, seed 1): assembly + blocked
        Cholesky + 2 triangular solves, whole;
      * cholesky_65536_sample: the N = 65536 construction (d = 16, b = 256,
        l = 2, sn2 = 1e-3, seed 3) at N = --cpu-n (14336 by default), the
        same steps, in TFLOP/s (N^3/3 over the time);
      * cg_32768_sample: the CG workload's system (N = 32768, b = 128,
        seed 2), assembled on the host, then --cpu-cg-iters CG iterations of
        the reference's CPU CG (matrixVectorBlock_CPU, chunked dot products),
        in GB/s on the GPU figure's basis (packed bytes per mat-vec: the
        iterations' mat-vecs, the %50 residual and the initial one).
    The headline `value` is the 65536-construction sample, in the bench
    line's unit."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from hs_testlib import Oracle

    syn = _load("hs_synth", os.path.join(PKG, "synthetic.py"))
    o = Oracle()
    cores = int(o.or_num_threads())
    per = {}

    def chol(seed, N, d, b, ell, sf2, sn2):
        X, y = syn.gp_inputs(seed, N, d)
        t0 = time.time()
        A = o.rbf(X, N, d, b, ell, sf2, sn2)
        L = o.cholesky(A, N, b)
        n = -(-N // b)
        rhs = np.zeros(n * b)
        rhs[:N] = y
        o.tri_solve(L, rhs, N, b)
        return time.time() - t0

    seed, N, d, b, ell, sf2, sn2 = syn.WORKLOADS["cholesky_small"]
    dt = chol(seed, N, d, b, ell, sf2, sn2)
    per["cholesky_4096"] = {"value": N ** 3 / 3.0 / dt / 1e12, "unit": "TFLOP/s", "seconds": dt,
                            "sample": "whole workload: assembly + Cholesky + 2 solves, N=%d b=%d d=%d" % (N, b, d)}
    log("cpu baseline: N=%d Cholesky workload %.2f s" % (N, dt))
    seed, _, d, b, ell, sf2, sn2 = syn.WORKLOADS["cholesky"]
    N = args.cpu_n
    dt = chol(seed, N, d, b, ell, sf2, sn2)
    per["cholesky_65536_sample"] = {
        "value": N ** 3 / 3.0 / dt / 1e12, "unit": "TFLOP/s", "seconds": dt,
        "sample": "assembly + Cholesky + 2 solves of the N=65536 construction at N=%d, b=%d, d=%d" % (N, b, d)}
    log("cpu baseline: N=%d sample of the 65536 construction %.2f s" % (N, dt))
    seed, N, d, b, ell, sf2, sn2 = syn.WORKLOADS["cg"]
    X, y = syn.gp_inputs(seed, N, d)
    A = o.rbf(X, N, d, b, ell, sf2, sn2)
    n = -(-N // b)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    iters = args.cpu_cg_iters
    t0 = time.time()
    _, it, _ = o.cg(A, rhs, N, b, 1e-300, imax=iters, chunks=cores)
    dt = time.time() - t0
    nsymv = it + -(-it // 50) + 1
    per["cg_32768_sample"] = {
        "value": nsymv * packed_len(N, b) * 8 / dt / 1e9, "unit": "GB/s", "seconds": dt, "iterations": it,
        "sample": "%d CG iterations (%d mat-vecs) of the N=%d, b=%d CG workload; %.2f s per iteration"
                  % (it, nsymv, N, b, dt / it)}
    log("cpu baseline: %d CG iterations at N=%d in %.2f s" % (it, N, dt))
    del A
    head = per["cholesky_65536_sample"]
    return {"value": head["value"], "unit": "TFLOP/s", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
            "host_cpus": os.cpu_count(), "sample": "oracle (C/OpenMP restatement of the reference's CPU path): " +
            head["sample"] + ", %.1f s" % head["seconds"], "per_workload": per}


BF16_PEAK_TFLOPS = 2500.0
FP16_PEAK_TFLOPS = 2500.0  # dense FP16 MFMA, the same rate as BF16
FP32_PEAK_TFLOPS = 157.3


def kernel_fractions(ch, mixed, cg, adj):
    """Achieved rate against the bounding MI355X peak for each hot kernel of
    the run (algorithmic flops or bytes over measured kernel time)."""
    out = {}
    up = ch["update"]
    out["trailing_mfma_f64 (FP64 SYRK+GEMM update)"] = {
        "achieved": up["tflops"], "unit": "TFLOP/s", "peak": FP64_PEAK_TFLOPS, "frac": up["tflops"] / FP64_PEAK_TFLOPS}
    out["FP64 factorisation, whole"] = {
        "achieved": ch["tflops_factor"], "unit": "TFLOP/s", "peak": FP64_PEAK_TFLOPS,
        "frac": ch["tflops_factor"] / FP64_PEAK_TFLOPS}
    n = -(-ch["N"] // ch["b"])
    packed = n * (n + 1) // 2 * ch["b"] ** 2 * 8
    asm = packed / (ch["assembly_ms"] / 1e3) / 1e9
    out["rbf assembly (K written)"] = {"achieved": asm, "unit": "GB/s", "peak": HBM_PEAK_GBS,
                                       "frac": asm / HBM_PEAK_GBS}
    sol = 2 * packed / (ch["solve_ms"] / 1e3) / 1e9
    out["triangular solves (L read twice, with the inverse build)"] = {
        "achieved": sol, "unit": "GB/s", "peak": HBM_PEAK_GBS, "frac": sol / HBM_PEAK_GBS}
    if mixed:
        f32 = mixed["fp32"]["update_tflops"]
        out["trailing_f32_ps256 + trailing_mfma_f32 (FP32 update)"] = {
            "achieved": f32, "unit": "TFLOP/s", "peak": FP32_PEAK_TFLOPS, "frac": f32 / FP32_PEAK_TFLOPS}
        b3 = mixed["bf16x3"]["update_tflops"]
        out["trailing_bf16x3_ps256 + trailing_mfma_bf16x3 (BF16 hi+lo update)"] = {
            "achieved": b3, "achieved_three_products": 3 * b3, "unit": "TFLOP/s", "peak": BF16_PEAK_TFLOPS,
            "frac": b3 / BF16_PEAK_TFLOPS, "frac_three_products": 3 * b3 / BF16_PEAK_TFLOPS}
        h3 = mixed["f16x3"]["update_tflops"]
        out["trailing_f16x3_ps256 + trailing_mfma_f16x3 (scaled FP16 hi+lo update)"] = {
            "achieved": h3, "achieved_three_products": 3 * h3, "unit": "TFLOP/s", "peak": FP16_PEAK_TFLOPS,
            "frac": h3 / FP16_PEAK_TFLOPS, "frac_three_products": 3 * h3 / FP16_PEAK_TFLOPS}
    if cg:
        out["symv_partials (read-once SYMV block pass)"] = {
            "achieved": cg["symv_block_pass_gbs"], "unit": "GB/s", "peak": HBM_PEAK_GBS,
            "frac": cg["symv_block_pass_gbs"] / HBM_PEAK_GBS, "stream_read_gbs": cg["stream_read_gbs"]}
        out["symv_ordered (reference-order mat-vec, parity mode)"] = {
            "achieved": cg["ref_order_symv_gbs"], "unit": "GB/s", "peak": HBM_PEAK_GBS,
            "frac": cg["ref_order_symv_gbs"] / HBM_PEAK_GBS}
    if adj:
        out["GP NLL + gradient (forward + reverse pass)"] = {
            "achieved": adj["tflops"], "unit": "TFLOP/s", "peak": FP64_PEAK_TFLOPS,
            "frac": adj["tflops"] / FP64_PEAK_TFLOPS}
    return out


def cpu_model():
    """The host CPU's model name from /proc/cpuinfo (None where it is not readable)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=0, help="override N of the Cholesky workload (testing only)")
    ap.add_argument("--no-cg", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=14336)
    ap.add_argument("--cpu-cg-iters", type=int, default=30)
    ap.add_argument("--no-gpr", action="store_true")
    ap.add_argument("--no-ref-order", action="store_true",
                    help="skip the untimed reference-order (parity mode) solve of the Cholesky workload")
    ap.add_argument("--only", choices=["cg", "symv", "adjoint", "gpr", "mixed"],
                    help="run only this side workload and print its record (for profiling; not a bench line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a kernel of known HBM traffic (y = 2x over 2^27 doubles, 5 launches) to scale "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.copy_calibration:
        hs = _load("hs_amd", os.path.join(PKG, "hs_amd.py"))
        with hs.Device(0) as dev:
            n = 1 << 27
            x, y = dev.zeros(n), dev.zeros(n)
            for _ in range(5):
                hs.call("hs_vec_scale_f64", dev.h, x.p, 2.0, y.p, 256, 0, n // 256)
            dev.sync()
        print(json.dumps({"calibration": "hs_vec_scale_f64", "launches": 5, "read_bytes": n * 8, "write_bytes": n * 8}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (no collectives, no sharding); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    hs = _load("hs_amd", os.path.join(PKG, "hs_amd.py"))
    syn = _load("hs_synth", os.path.join(PKG, "synthetic.py"))
    chol_n = args.N or syn.WORKLOADS["cholesky"][1]
    # every system the run solves, reserved up front: no solver entry allocates
    sizes = [(chol_n, syn.WORKLOADS["cholesky"][3]), (syn.WORKLOADS["cg"][1], syn.WORKLOADS["cg"][3], 8192),
             (syn.WORKLOADS["adjoint"][1], syn.WORKLOADS["adjoint"][3])]
    if args.only:
        with hs.Device(0, reserve=sizes) as dev:
            if args.only == "adjoint":
                rec = adjoint_workload(hs, dev, syn, args)
            elif args.only == "gpr":
                rec = gpr_workload(hs, dev, syn, args)
            elif args.only == "mixed":
                rec = mixed_workload(hs, dev, syn, args)
            else:
                rec = cg_workload(hs, dev, syn, args, symv_only=(args.only == "symv"))
        print(json.dumps({args.only: rec}))
        return
    with hs.Device(0, reserve=sizes) as dev:
        ch = cholesky_workload(hs, dev, syn, args)
        x64 = ch.pop("_x")
        mixed = None if args.no_mixed else mixed_workload(hs, dev, syn, args, x64, ch["rel_residual"])
        cg = None if args.no_cg else cg_workload(hs, dev, syn, args)
        adj = None if args.no_adjoint else adjoint_workload(hs, dev, syn, args)
        gpr = None if args.no_gpr else gpr_workload(hs, dev, syn, args)
    if mixed:
        fp64_solve = ch["factor_ms"] + ch["solve_ms"]
        for v in mixed.values():
            for t in ("backward_error_1e-12", "equal_residual"):
                if t in v:
                    v[t]["speedup_vs_fp64_factor_plus_solve"] = fp64_solve / v[t]["solve_ms"]
    cpu = None if args.no_cpu else cpu_baseline(args)
    up = ch["update"]
    out = {
        "metric": METRIC,
        "value": ch["tflops_step"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ch["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic GP inputs (SplitMix64 seed %d, X ~ U[-1,1]^{N x d})" % ch["seed"],
        "config": {"workload": "FP64 RBF assembly + blocked Cholesky + 2 triangular solves",
                   "N": ch["N"], "matrix_bsz": ch["b"], "d": ch["d"], "lengthscale": ch["ell"],
                   "signal_variance": ch["sf2"], "noise_variance": ch["sn2"], "parallelism": "1 GPU"},
        "roofline": {"kernel": "trailing_mfma_f64 (fused SYRK+GEMM trailing update)", "bound": "mfma",
                     "achieved": up["tflops"], "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": up["tflops"] / FP64_PEAK_TFLOPS,
                     "traffic": (TRAILING_TRAFFIC_BYTES if (ch["N"], ch["b"], up["launches"]) == (65536, 256, 62)
                                 else None),
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)",
                     "flops_per_launch": up["flops_per_launch"], "avg_launch_ms": up["avg_launch_ms"],
                     "launches": up["launches"]},
        "cpu_baseline": cpu,
        "cholesky": ch,
        "mixed": mixed,
        "cg": cg,
        "adjoint": adj,
        "gpr": gpr,
        "kernels": kernel_fractions(ch, mixed, cg, adj),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
