"""Benchmark: triangle-soup fast winding number on one MI355X.

Workload (the headline configuration): level-6 icosphere, 81,920 triangles,
radially displaced by 16 seeded Gaussian bumps; 1,000,000 queries, half
uniform in the bbox and half within 1e-3 of the surface; order 2.

A step is one query pass over the 1M-query batch (Morton ordering of the
queries + traversal), with the tree built beforehand and the queries resident
in HBM.  accuracy_scale (beta) is chosen as the smallest value on a grid for
which max |W - W_exact| <= 1e-4 over all 1M queries, W_exact being the FP64
exact sum (GPU FP64 kernel, itself checked against the CPU FP64 oracle on a
subsample here).  The CPU baseline is the reference's own UT_SolidAngle
(compiled from the reference sources into oracle/_ref) run with
igl::parallel_for on the host cores at the same beta.

Prints ONE JSON line (rank 0).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "winding queries/s at max |dw|<=1e-4 vs FP64 oracle"
ACC_TARGET = 1e-4
BETA_GRID = [2.0, 3.0, 3.5, 4.0, 4.25, 4.5, 4.6, 4.7, 4.8, 4.9, 5.0, 5.25, 5.5, 6.0, 7.0, 8.0]
# Counted FLOPs per traversal event (add/mul/div/sqrt = 1, atan2 = 1):
FLOP_TEST = 10    # |q-P|^2 <= maxP2 * beta^2 per child slot (7179-7191)
FLOP_FAR = 83     # accepted order-2 far-field slot (7202-7241)
FLOP_EXACT = 76   # exact triangle (6078-6109)
# HBM bytes moved per k_traverse launch on the default workload (1M queries,
# SAH4 tree), keyed by accuracy_scale: rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (KiB per dispatch), each scaled by what the same counter reads
# for a 256 MiB device copy (fetch x2.000, write x1.000).
MEASURED_TRAFFIC = {4.9: int((181.0 + 33.4) * 2**20)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nq", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--beta", type=float, default=0.0,
                    help="skip the accuracy sweep and time this accuracy_scale only (profiling runs)")
    ap.add_argument("--tree", choices=["sah4", "lbvh"], default="sah4")
    ap.add_argument("--no-mixed", action="store_true", help="skip the 1.3M-triangle mixed vs FP64 comparison")
    ap.add_argument("--no-rows", action="store_true", help="skip the cloud / exact / signed-distance legs")
    ap.add_argument("--rows-only", action="store_true",
                    help="run only the rows named by --rows and print them as one JSON line (profiling runs)")
    ap.add_argument("--rows", default="cloud,exact,signed_distance,point_areas,antipodal,curve",
                    help="comma-separated rows reported beside the headline")
    ap.add_argument("--serving-duration", type=float, default=2.0, help="seconds of Poisson trace (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run 3 device-to-device copies of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"mode": "copy", "bytes": 4 * n, "copies": 3}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (more GPUs would shard the "
                                   "queries by curve-order range against a replicated tree)"}))
        sys.exit(2)

    import numpy as np
    import torch

    from libigl_amd import datagen as D
    from libigl_amd import gwn
    from libigl_amd import _native as nat

    if args.rows_only:
        torch.cuda.synchronize()
        print(json.dumps({"rows": run_rows(args.rows, args.no_cpu)}), flush=True)
        return

    dev = torch.device("cuda:0")
    t0 = time.time()
    V, F, Q = D.soup_workload(nq=args.nq)
    log(f"[bench] workload: {F.shape[0]} tris, {Q.shape[0]} queries ({time.time() - t0:.1f}s)")

    bvh = gwn.FastWindingNumberBVH()
    torch.cuda.synchronize()
    tb = time.perf_counter()
    tree = gwn.TREE_SAH4 if args.tree == "sah4" else gwn.TREE_LBVH
    bvh.build(V, F, 2, tree)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - tb
    # the first build of the process also loads the library's device code;
    # a second build of the same mesh times the build alone
    tb = time.perf_counter()
    gwn.FastWindingNumberBVH().build(V, F, 2, tree)
    torch.cuda.synchronize()
    build_warm_s = time.perf_counter() - tb
    info = bvh.info()
    log(f"[bench] {args.tree.upper()} tree: {info.num_nodes} nodes, {info.num_slots} slots, depth {info.depth}, "
        f"build {build_s:.3f}s (first in process), {build_warm_s:.4f}s (again)")

    nq = Q.shape[0]
    Qf = Q.astype(np.float32)
    dQ = torch.from_numpy(Qf).to(dev)
    dW = torch.empty(nq, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)

    def run(beta, prec=gwn.PREC_FP32):
        bvh.query_device(dQ.data_ptr(), nq, dW.data_ptr(), beta, prec, stream.cuda_stream)

    if args.beta > 0:  # profiling mode: only the timed steps at one beta
        for _ in range(args.warmup):
            run(args.beta)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            run(args.beta)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        print(json.dumps({"mode": "profile", "beta": args.beta, "tree": args.tree, "ms_per_step": 1e3 * dt,
                          "queries_per_s": nq / dt}), flush=True)
        return

    # FP64 exact reference on the GPU for all queries
    # the oracle sees the original FP64 mesh and queries (the FWN path narrows both to float)
    dV = torch.from_numpy(np.ascontiguousarray(V, dtype=np.float64)).to(dev)
    dF = torch.from_numpy(F).to(dev)
    dQ64 = torch.from_numpy(np.ascontiguousarray(Q, dtype=np.float64)).to(dev)
    dWx = torch.empty(nq, dtype=torch.float64, device=dev)
    te = time.perf_counter()
    nat.check(nat.lib().gwn_exact_winding_number_device(dV.data_ptr(), V.shape[0], dF.data_ptr(), F.shape[0],
                                                       dQ64.data_ptr(), nq, dWx.data_ptr(), stream.cuda_stream))
    torch.cuda.synchronize()
    exact_s = time.perf_counter() - te
    Wx = dWx.cpu().numpy()
    log(f"[bench] FP64 exact (GPU) for all queries: {exact_s:.2f}s")

    # confirm the GPU exact sum against the CPU FP64 oracle on a subsample
    oracle_check = None
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        sub = np.arange(0, nq, max(1, nq // 2000))
        Wo = O.winding_number_bf(V, F, Q[sub])
        oracle_check = float(np.abs(Wo - Wx[sub]).max())
        log(f"[bench] GPU exact vs CPU FP64 oracle on {sub.size} queries: max {oracle_check:.2e}")
    except Exception as e:  # the oracle is optional for the measurement itself
        log(f"[bench] oracle check skipped: {e}")

    # choose beta: smallest grid value meeting the accuracy target on all queries
    err_at = {}
    beta_acc = None
    for beta in BETA_GRID:
        run(beta)
        torch.cuda.synchronize()
        e = np.abs(dW.cpu().numpy() - Wx)
        err_at[beta] = (float(e.max()), float(np.quantile(e, 0.99)))
        log(f"[bench] beta {beta}: max|dw| {err_at[beta][0]:.3e} p99 {err_at[beta][1]:.3e}")
        if err_at[beta][0] <= ACC_TARGET:
            beta_acc = beta
            break
    if beta_acc is None:
        beta_acc = BETA_GRID[-1]

    def timed(beta, steps, warmup, prec=gwn.PREC_FP32):
        for _ in range(warmup):
            run(beta, prec)
        torch.cuda.synchronize()
        nat.lib().gwn_profile_reset()
        nat.lib().gwn_profile_enable(1)
        t = time.perf_counter()
        for _ in range(steps):
            run(beta, prec)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        nat.lib().gwn_profile_enable(0)
        import ctypes
        ms, n = ctypes.c_double(), ctypes.c_int64()
        name = b"soup_traverse64" if prec == gwn.PREC_FP64 else b"soup_traverse"
        nat.lib().gwn_profile_read(name, ctypes.byref(ms), ctypes.byref(n))
        ks, kn = ctypes.c_double(), ctypes.c_int64()
        nat.lib().gwn_profile_read(b"radix_sort", ctypes.byref(ks), ctypes.byref(kn))
        return dt, ms.value / max(n.value, 1), ks.value / max(kn.value, 1)

    dt, k_ms, sort_ms = timed(beta_acc, args.steps, args.warmup)
    ms_step = 1000 * dt / args.steps
    qps = nq / (dt / args.steps)
    log(f"[bench] beta {beta_acc}: {ms_step:.3f} ms/step, {qps:.3e} q/s; traverse {k_ms:.3f} ms, sort {sort_ms:.3f} ms")
    dt2, k2_ms, _ = timed(2.0, max(3, args.steps // 2), 1)
    qps_b2 = nq / (dt2 / max(3, args.steps // 2))
    dtm, km_ms, _ = timed(beta_acc, max(3, args.steps // 2), 1, gwn.PREC_MIXED)
    dtd, kd_ms, _ = timed(beta_acc, max(3, args.steps // 4), 1, gwn.PREC_FP64)
    run(beta_acc, gwn.PREC_MIXED)
    torch.cuda.synchronize()
    err_mixed = float(np.abs(dW.cpu().numpy() - Wx).max())
    run(beta_acc, gwn.PREC_FP64)
    torch.cuda.synchronize()
    err_fp64 = float(np.abs(dW.cpu().numpy() - Wx).max())

    # counted FLOPs per query from the oracle's instrumented traversal of the
    # same tree on a sample of the queries
    flops_per_q = None
    try:
        soup_o = O.Soup(V, F)
        samp = np.arange(0, nq, max(1, nq // 4000))
        tot = np.zeros(3)
        for i in samp:
            v, a, x = soup_o.count(Qf[i], beta_acc)
            tot += (v, a, x)
        visits, acc, exact = tot / samp.size
        flops_per_q = FLOP_TEST * 4 * visits + FLOP_FAR * acc + FLOP_EXACT * exact
        log(f"[bench] per query: {visits:.1f} node visits, {acc:.1f} far slots, {exact:.1f} exact tris "
            f"-> {flops_per_q:.0f} FLOP")
    except Exception as e:
        log(f"[bench] flop count skipped: {e}")

    roofline = None
    if flops_per_q is not None and k_ms > 0:
        achieved = flops_per_q * nq / (k_ms * 1e-3) / 1e12
        roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": 157.3, "unit": "TFLOP/s",
                    "frac": round(achieved / 157.3, 4),
                    "traffic": (MEASURED_TRAFFIC.get(beta_acc)
                                if (nq == 1_000_000 and args.tree == "sah4") else None),
                    "hbm_GBps": (round(MEASURED_TRAFFIC[beta_acc] / (k_ms * 1e-3) / 1e9, 1)
                                 if (beta_acc in MEASURED_TRAFFIC and nq == 1_000_000 and args.tree == "sah4")
                                 else None),
                    "hbm_peak_GBps": 8000.0,
                    "kernel": "k_traverse<2,false>", "kernel_ms": round(k_ms, 4),
                    "flop_per_query": round(flops_per_q, 1),
                    "note": "FP32 VALU; no MFMA on this path; packed FP32 disabled so 78.6 TFLOP/s is the "
                            "attainable scalar-FP32 ceiling"}

    cpu = None
    if not args.no_cpu:
        try:
            threads = args.cpu_threads or min(16, os.cpu_count() or 1)
            os.environ["IGL_NUM_THREADS"] = str(threads)
            if O.ref_available():
                ref = O.RefSoup(V, F)
                kind = "reference"
            else:
                ref = O.Soup(V, F)
                kind = "port"
            csub = min(nq, 1_000_000)
            Qc = Q[:csub]
            runs = []
            for rep in range(3):
                tc = time.perf_counter()
                if kind == "reference":
                    ref.query(Qc, beta_acc)
                else:
                    ref.query(Qc, beta_acc, threads)
                runs.append(time.perf_counter() - tc)
                if sum(runs) > 20:
                    break
            cdt = float(np.median(runs))
            nthr = O.ref().ref_num_threads() if kind == "reference" else threads
            cpu = {"value": round(csub / cdt, 1), "unit": "queries/s", "cores": int(nthr), "kind": kind,
                   "sample": f"{csub} queries of the same workload at beta={beta_acc}, median of {len(runs)}; "
                             f"UT_SolidAngle + igl::parallel_for"}
            log(f"[bench] CPU ({kind}, {nthr} threads): {csub / cdt:.3e} q/s")
        except Exception as e:
            log(f"[bench] cpu baseline skipped: {e}")

    mixed_1m_tri = None
    if not args.no_mixed:
        try:
            mixed_1m_tri = run_mixed_workload(BETA_GRID)
        except Exception as e:
            log(f"[bench] mixed workload skipped: {e}")

    others = {} if args.no_rows else run_rows(args.rows, args.no_cpu)

    serving = None
    if args.serving_duration > 0:
        try:
            serving = run_serving(args.serving_duration, beta_acc)
        except Exception as e:
            log(f"[bench] serving skipped: {e}")

    out = {
        "metric": METRIC, "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded generator)",
        "config": {"workload": "soup fast_winding_number, level-6 icosphere 81,920 tris, 16 bumps, 1M queries "
                               "(50% uniform, 50% within 1e-3 of surface), order 2",
                   "accuracy_scale": beta_acc, "tree": "SAH4 (reference topology)", "parallelism": "single GPU"},
        "accuracy": {"max_abs_dw_vs_fp64": err_at[beta_acc][0], "p99_abs_dw_vs_fp64": err_at[beta_acc][1],
                     "gpu_exact_vs_cpu_oracle": oracle_check, "by_beta": {str(k): v for k, v in err_at.items()}},
        "beta2_queries_per_s": round(qps_b2, 1),
        "mixed": {"queries_per_s": round(nq / (dtm / max(3, args.steps // 2)), 1), "max_abs_dw": err_mixed,
                  "kernel_ms": round(km_ms, 4)},
        "fp64": {"queries_per_s": round(nq / (dtd / max(3, args.steps // 4)), 1), "max_abs_dw": err_fp64,
                 "kernel_ms": round(kd_ms, 4)},
        "build_s": round(build_s, 4), "build_warm_s": round(build_warm_s, 4), "sort_ms": round(sort_ms, 4),
        "roofline": roofline, "cpu_baseline": cpu, "serving": serving, "mixed_vs_fp64": mixed_1m_tri,
        "other_rows": others,
    }
    print(json.dumps(out), flush=True)


def run_rows(rows, no_cpu):
    """The rows reported beside the headline, by name (comma-separated)."""
    wanted = set(rows.split(","))
    out = {}
    for name, fn in (("cloud", run_cloud), ("exact", run_exact), ("signed_distance", run_signed_distance),
                     ("point_areas", run_point_areas), ("antipodal", run_antipodal), ("curve", run_curve)):
        if name not in wanted:
            continue
        try:
            out[name] = fn(no_cpu)
        except Exception as e:
            log(f"[bench] {name} skipped: {e}")
    return out


def run_mixed_workload(beta_grid):
    """Mixed-precision vs all-FP64 soup queries on the level-8 icosphere
    (1,310,720 triangles, 64 bumps), 10M queries (80% uniform, 20% within 1e-4
    of the surface).  |dw| against the FP64 exact sum on a 100k-query sample."""
    import numpy as np
    import torch

    from libigl_amd import datagen as D
    from libigl_amd import gwn
    from libigl_amd import _native as nat

    t0 = time.time()
    V, F, Q = D.mixed_workload()
    nq = Q.shape[0]
    bvh = gwn.FastWindingNumberBVH()
    tb = time.perf_counter()
    bvh.build(V, F, 2, gwn.TREE_SAH4)
    build_s = time.perf_counter() - tb
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream(dev)
    dQ = torch.from_numpy(Q.astype(np.float32)).to(dev)
    dW = torch.empty(nq, dtype=torch.float64, device=dev)
    sub = np.arange(0, nq, nq // 100_000)
    dV = torch.from_numpy(np.ascontiguousarray(V, dtype=np.float64)).to(dev)
    dF = torch.from_numpy(F).to(dev)
    dQs = torch.from_numpy(np.ascontiguousarray(Q[sub], dtype=np.float64)).to(dev)
    dWx = torch.empty(sub.size, dtype=torch.float64, device=dev)
    nat.check(nat.lib().gwn_exact_winding_number_device(dV.data_ptr(), V.shape[0], dF.data_ptr(), F.shape[0],
                                                       dQs.data_ptr(), sub.size, dWx.data_ptr(), st.cuda_stream))
    torch.cuda.synchronize()
    Wx = dWx.cpu().numpy()
    log(f"[bench] mixed workload: {F.shape[0]} tris, {nq} queries, build {build_s:.2f}s, setup {time.time() - t0:.1f}s")

    def run(beta, prec):
        bvh.query_device(dQ.data_ptr(), nq, dW.data_ptr(), beta, prec, st.cuda_stream)

    def err(beta, prec):
        run(beta, prec)
        torch.cuda.synchronize()
        e = np.abs(dW.cpu().numpy()[sub] - Wx)
        return float(e.max()), float(np.quantile(e, 0.99))

    beta = beta_grid[-1]
    for b in beta_grid:
        if err(b, gwn.PREC_MIXED)[0] <= ACC_TARGET:
            beta = b
            break

    def rate(prec, steps):
        run(beta, prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            run(beta, prec)
        torch.cuda.synchronize()
        return nq * steps / (time.perf_counter() - t)

    out = {"workload": "level-8 icosphere 1,310,720 tris, 64 bumps, 10M queries (80% uniform, 20% within 1e-4)",
           "accuracy_scale": beta, "build_s": round(build_s, 3), "error_sample": int(sub.size)}
    for name, prec in (("fp32", gwn.PREC_FP32), ("mixed", gwn.PREC_MIXED), ("fp64", gwn.PREC_FP64)):
        mx, p99 = err(beta, prec)
        out[name] = {"queries_per_s": round(rate(prec, 3), 1), "max_abs_dw": mx, "p99_abs_dw": p99}
        log(f"[bench] 1M-triangle mesh, 10M queries, {name}: {out[name]}")
    out["mixed_over_fp64"] = round(out["mixed"]["queries_per_s"] / out["fp64"]["queries_per_s"], 3)
    del dQ, dW, dV, dF, dQs, dWx
    return out


def _kernel_ms(name):
    import ctypes

    from libigl_amd import _native as nat
    ms, n = ctypes.c_double(), ctypes.c_int64()
    nat.lib().gwn_profile_read(name.encode(), ctypes.byref(ms), ctypes.byref(n))
    return ms.value / max(n.value, 1)


def _timed_device(fn, steps, kernel):
    """warm-up call, then `steps` calls between synchronizations; returns
    (seconds per call, average kernel ms from HIP events)."""
    import torch

    from libigl_amd import _native as nat
    fn()
    torch.cuda.synchronize()
    nat.lib().gwn_profile_reset()
    nat.lib().gwn_profile_enable(1)
    t = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / steps
    nat.lib().gwn_profile_enable(0)
    return dt, _kernel_ms(kernel)


def run_cloud(no_cpu):
    """Point-cloud path: 2M oriented points on the noisy level-5 sphere, 160^3
    grid queries, order 2, beta 2; W checked against the oracle on a sample."""
    import numpy as np
    import torch

    from libigl_amd import _native as nat
    from libigl_amd import datagen as D
    from libigl_amd import gwn

    P, N, A, Q = D.cloud_workload()
    tb = time.perf_counter()
    c = gwn.PointCloudFWN(P, N, A, 2)
    build_s = time.perf_counter() - tb
    tb = time.perf_counter()
    c = gwn.PointCloudFWN(P, N, A, 2)  # again, now that the device code is loaded
    build_warm_s = time.perf_counter() - tb
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream(dev).cuda_stream
    nq = Q.shape[0]
    dQ = torch.from_numpy(Q.astype(np.float32)).to(dev)
    dW = torch.empty(nq, dtype=torch.float32, device=dev)
    dt, kms = _timed_device(lambda: nat.check(nat.lib().gwn_cloud_query_device(c._h, dQ.data_ptr(), nq, 2.0,
                                                                               dW.data_ptr(), st)),
                            5, "cloud_traverse")
    W = dW.cpu().numpy()
    out = {"workload": "2M points, 160^3 grid queries, order 2, beta 2", "cells": int(c.ncells()),
           "build_s": round(build_s, 4), "build_warm_s": round(build_warm_s, 4), "queries_per_s": round(nq / dt, 1),
           "kernel_ms": round(kms, 4)}
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        oc = O.Cloud(P.astype(np.float32), N.astype(np.float32), A.astype(np.float32), 2, np.float32)
        sub = np.arange(0, nq, nq // 20_000)
        threads = min(16, os.cpu_count() or 1)
        t = time.perf_counter()
        Wo = oc.eval(Q[sub].astype(np.float32), 2.0, threads)
        cdt = time.perf_counter() - t
        out["max_abs_dw_vs_oracle"] = float(np.abs(W[sub] - Wo).max())
        out["cpu_baseline"] = {"value": round(sub.size / cdt, 1), "unit": "queries/s", "cores": threads,
                               "kind": "port", "sample": f"{sub.size} grid queries, oracle restatement"}
    log(f"[bench] cloud: {out}")
    return out


def run_exact(no_cpu):
    """Exact FP64 winding number: level-5 noisy icosphere (20,480 triangles),
    100k queries uniform in [-1.5,1.5]^3."""
    import numpy as np
    import torch

    from libigl_amd import _native as nat
    from libigl_amd import datagen as D

    V, F, Q = D.exact_workload()
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream(dev).cuda_stream
    dV = torch.from_numpy(np.ascontiguousarray(V)).to(dev)
    dF = torch.from_numpy(F).to(dev)
    dQ = torch.from_numpy(np.ascontiguousarray(Q)).to(dev)
    dW = torch.empty(Q.shape[0], dtype=torch.float64, device=dev)
    dt, kms = _timed_device(lambda: nat.check(nat.lib().gwn_exact_winding_number_device(
        dV.data_ptr(), V.shape[0], dF.data_ptr(), F.shape[0], dQ.data_ptr(), Q.shape[0], dW.data_ptr(), st)),
        5, "exact")
    W = dW.cpu().numpy()
    pairs = Q.shape[0] * F.shape[0]
    out = {"workload": "20,480 tris, 100k queries, FP64", "queries_per_s": round(Q.shape[0] / dt, 1),
           "kernel_ms": round(kms, 4), "pairs_per_s": round(pairs / (kms * 1e-3), 1)}
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        sub = np.arange(0, Q.shape[0], 50)
        threads = min(16, os.cpu_count() or 1)
        t = time.perf_counter()
        Wo = O.winding_number_bf(V, F, Q[sub], threads)
        cdt = time.perf_counter() - t
        out["max_abs_dw_vs_oracle"] = float(np.abs(W[sub] - Wo).max())
        out["cpu_baseline"] = {"value": round(sub.size / cdt, 1), "unit": "queries/s", "cores": threads,
                               "kind": "port", "sample": f"{sub.size} queries, brute-force FP64 restatement"}
    log(f"[bench] exact: {out}")
    return out


def run_signed_distance(no_cpu):
    """igl::signed_distance with the fast-winding-number sign on the 81,920-
    triangle workload mesh, 1M points (closest point + FWN sign)."""
    import numpy as np
    import torch

    from libigl_amd import _native as nat
    from libigl_amd import datagen as D
    from libigl_amd import gwn

    V, F, Q = D.soup_workload()
    bvh = gwn.FastWindingNumberBVH().build(V, F, 2)
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream(dev).cuda_stream
    n = Q.shape[0]
    dP = torch.from_numpy(np.ascontiguousarray(Q)).to(dev)
    dS = torch.empty(n, dtype=torch.float64, device=dev)
    dI = torch.empty(n, dtype=torch.int32, device=dev)
    dC = torch.empty((n, 3), dtype=torch.float64, device=dev)
    lo, hi = gwn.DBL_MIN, gwn.DBL_MAX
    dt, kms = _timed_device(lambda: nat.check(nat.lib().gwn_signed_distance_device(
        bvh.handle, dP.data_ptr(), n, gwn.SIGN_FAST_WINDING_NUMBER, lo, hi, dS.data_ptr(), dI.data_ptr(),
        dC.data_ptr(), st)), 3, "closest_point")
    S = dS.cpu().numpy()
    out = {"workload": "81,920 tris, 1M points, FAST_WINDING_NUMBER sign", "points_per_s": round(n / dt, 1),
           "closest_point_kernel_ms": round(kms, 4)}
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        sub = np.arange(0, n, 2500)
        So, _, _ = O.signed_distance(V, F, Q[sub], gwn.SIGN_FAST_WINDING_NUMBER)
        out["max_rel_dS_vs_oracle"] = float(np.max(np.abs(S[sub] - So) / np.maximum(np.abs(So), 1e-12)))
    log(f"[bench] signed distance: {out}")
    return out


def run_point_areas(no_cpu):
    """Areas for a cloud that has none (copyleft/cgal/fast_winding_number.cpp:
    32-43) on the 2M oriented points of the point-cloud workload: knn with
    k = 21 over all points, point areas, and the whole estimated-area cloud
    build (coincident-point filter, knn, areas, octree, moments)."""
    import numpy as np
    import torch

    from libigl_amd import _native as nat
    from libigl_amd import datagen as D
    from libigl_amd import gwn

    P, N, _, _ = D.cloud_workload()
    n, k = P.shape[0], 21
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream(dev).cuda_stream
    dP = torch.from_numpy(np.ascontiguousarray(P)).to(dev)
    dN = torch.from_numpy(np.ascontiguousarray(N)).to(dev)
    dI = torch.empty((n, k), dtype=torch.int32, device=dev)
    dA = torch.empty(n, dtype=torch.float64, device=dev)
    dT = torch.empty((n, 3), dtype=torch.float64, device=dev)
    dt_knn, kms_knn = _timed_device(lambda: nat.check(nat.lib().gwn_knn_device(
        dP.data_ptr(), n, dP.data_ptr(), n, k, None, 0, dI.data_ptr(), st)), 3, "knn")
    dt_pa, kms_pa = _timed_device(lambda: nat.check(nat.lib().gwn_point_areas_device(
        dP.data_ptr(), n, dI.data_ptr(), k, dN.data_ptr(), dA.data_ptr(), dT.data_ptr(), st)), 3, "point_areas")
    tb = time.perf_counter()
    c = gwn.PointCloudFWN.with_estimated_areas(P, N, 2)
    build_s = time.perf_counter() - tb
    I, A = dI.cpu().numpy(), dA.cpu().numpy()
    out = {"workload": "2M oriented surface points (noisy level-5 sphere), k = 21",
           "knn_points_per_s": round(n / dt_knn, 1), "knn_kernel_ms": round(kms_knn, 4),
           "point_areas_points_per_s": round(n / dt_pa, 1), "point_areas_kernel_ms": round(kms_pa, 4),
           "estimated_area_cloud_build_s": round(build_s, 4),
           "area_sum_over_surface_area": None}
    Vs, Fs = D.noisy_sphere(5, 0.01, 11)
    out["area_sum_over_surface_area"] = round(float(A.sum() / D.face_normals(Vs, Fs)[1].sum()), 6)
    assert np.array_equal(c.A, A)
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        threads = min(16, os.cpu_count() or 1)
        tree = O.Octree(P, np.float64)
        sub = np.arange(0, n, 50)
        t = time.perf_counter()
        Io = O.knn(P[sub], P, k, tree, threads)
        cdt = time.perf_counter() - t
        out["knn_matches_oracle"] = bool(np.array_equal(I[sub], Io))
        out["cpu_baseline"] = {"value": round(sub.size / cdt, 1), "unit": "points/s (knn, k = 21)",
                               "cores": threads, "kind": "port",
                               "sample": f"{sub.size} queries, best-first octree search restated (knn.cpp:72-127)"}
        rows = sub[:400]
        t = time.perf_counter()
        Ao, _ = O.point_areas(P, I, N, rows=rows)
        adt = time.perf_counter() - t
        out["point_areas_max_rel_vs_oracle"] = float(np.max(np.abs(A[rows] - Ao[rows]) / np.maximum(Ao[rows], 1e-300)))
        out["cpu_baseline_point_areas"] = {"value": round(rows.size / adt, 1), "unit": "points/s", "cores": 1,
                                           "kind": "port",
                                           "sample": f"{rows.size} points, numpy SVD + Qhull Delaunay restatement"}
    log(f"[bench] point areas: {out}")
    return out


def run_antipodal(no_cpu):
    """Exact winding numbers by the antipodal method (WindingNumberAntipodalScene):
    the 81,920-triangle workload mesh (closed: the boundary term is empty and W
    is the signed ray-crossing count), 1M queries; checked on a sample against
    the exact FP64 sum."""
    import numpy as np
    import torch

    from libigl_amd import _native as nat
    from libigl_amd import datagen as D
    from libigl_amd import gwn

    V, F, Q = D.soup_workload()
    tb = time.perf_counter()
    scene = gwn.WindingNumberAntipodalScene(V, F)
    build_s = time.perf_counter() - tb
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream(dev).cuda_stream
    n = Q.shape[0]
    dQ = torch.from_numpy(np.ascontiguousarray(Q)).to(dev)
    dW = torch.empty(n, dtype=torch.float64, device=dev)
    dt, kms = _timed_device(lambda: nat.check(nat.lib().gwn_antipodal_query_device(
        scene._h, dQ.data_ptr(), n, dW.data_ptr(), st)), 5, "antipodal_crossings")
    W = dW.cpu().numpy()
    out = {"workload": "81,920 tris (closed), 1M queries, FP64", "queries_per_s": round(n / dt, 1),
           "crossings_kernel_ms": round(kms, 4), "build_s": round(build_s, 4),
           "integer_valued": bool(np.all(W == np.round(W)))}
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        sub = np.arange(0, n, 1000)
        threads = min(16, os.cpu_count() or 1)
        We = O.winding_number_bf(V, F, Q[sub], threads)
        out["max_abs_dw_vs_exact"] = float(np.abs(W[sub] - We).max())
        t = time.perf_counter()
        Wo, _ = O.winding_number_antipodal(V, F, Q[sub], None, threads)
        cdt = time.perf_counter() - t
        out["matches_oracle"] = bool(np.array_equal(W[sub], Wo))
        out["cpu_baseline"] = {"value": round(sub.size / cdt, 1), "unit": "queries/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{sub.size} queries, restatement with brute-force crossings (the reference "
                                         "counts them with Embree, which is not in the image)"}
    log(f"[bench] antipodal: {out}")
    return out


# counted FLOPs of the 2D path (UT_SubtendedAngle, FastWindingNumberForSoups.h
# 7725-7809): the near/far test of one child slot (q - P, |q - P|^2, the
# scaled radius), an accepted order-2 far field, an exact segment (atan2 as 1)
FLOP_TEST2, FLOP_FAR2, FLOP_SEG = 6, 44, 12


def run_curve(no_cpu):
    """2D fast winding number (UT_SubtendedAngle<float,float>::computeAngle):
    a 262,144-segment bumpy closed curve, 1M queries (half uniform in its box,
    half within 1e-3 of the curve), order 2, accuracy_scale 2 (computeAngle's
    default); checked on a sample against the oracle (bit-exact with the
    reference) and against exact evaluation."""
    import numpy as np
    import torch

    from libigl_amd import _native as nat
    from libigl_amd import datagen as D
    from libigl_amd import gwn

    P, E, Q = D.curve_workload()
    n = Q.shape[0]
    torch.cuda.synchronize()
    tb = time.perf_counter()
    c = gwn.SubtendedAngle(P, E, 2)
    build_s = time.perf_counter() - tb
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream(dev).cuda_stream
    Qf = np.ascontiguousarray(Q, dtype=np.float32)
    dQ = torch.from_numpy(Qf).to(dev)
    dA = torch.empty(n, dtype=torch.float32, device=dev)
    dt, kms = _timed_device(lambda: nat.check(nat.lib().gwn_curve_query_device(
        c._h, dQ.data_ptr(), n, 2.0, dA.data_ptr(), st)), 20, "curve_traverse")
    A = dA.cpu().numpy()
    info = c.info()
    out = {"workload": "bumpy closed curve, 262,144 segments, 1M queries (50% uniform, 50% within 1e-3), "
                       "order 2, accuracy_scale 2, FP32", "queries_per_s": round(n / dt, 1),
           "kernel": "k_curve_traverse<2>", "kernel_ms": round(kms, 4), "build_s": round(build_s, 4),
           "tree": {"nodes": int(info.num_nodes), "slots": int(info.num_slots), "depth": int(info.depth)}}
    sub = np.arange(0, n, 1000)
    exact = c.compute_angle(Qf[sub], 1e6)  # every segment evaluated exactly
    out["max_abs_dw_vs_exact"] = float(np.abs(A[sub] - exact).max() / (2 * np.pi))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    co = O.Curve(P, E)
    tot = np.zeros(3)
    samp = np.arange(0, n, max(1, n // 4000))
    for i in samp:
        tot += co.count(Qf[i], 2.0)
    visits, acc, ex = tot / samp.size
    fpq = FLOP_TEST2 * 4 * visits + FLOP_FAR2 * acc + FLOP_SEG * ex
    if kms > 0:
        achieved = fpq * n / (kms * 1e-3) / 1e12
        out["roofline"] = {"bound": "valu", "achieved": round(achieved, 3), "peak": 157.3, "unit": "TFLOP/s",
                           "frac": round(achieved / 157.3, 4), "flop_per_query": round(fpq, 1),
                           "per_query": {"node_visits": round(visits, 2), "far_slots": round(acc, 2),
                                         "exact_segments": round(ex, 2)}}
    Ao = co.angle(Qf[sub], 2.0)
    out["max_abs_dw_vs_oracle"] = float(np.abs(A[sub].astype(np.float64) - Ao).max() / (2 * np.pi))
    # exact 2D winding number (igl::winding_number with F.cols() == 2) on the
    # same curve: 20,000 of the queries against all 262,144 edges, FP64
    Oe = np.ascontiguousarray(Q[::50][:20_000], dtype=np.float64)
    gwn.winding_number(P, E, Oe)
    nat.lib().gwn_profile_reset()
    nat.lib().gwn_profile_enable(1)
    te = time.perf_counter()
    for _ in range(3):
        We = gwn.winding_number(P, E, Oe)
    ete = (time.perf_counter() - te) / 3
    nat.lib().gwn_profile_enable(0)
    ekms = _kernel_ms("exact2d")
    ex2 = {"queries": int(Oe.shape[0]), "edges": int(E.shape[0]), "kernel_ms": round(ekms, 4),
           "pairs_per_s": round(Oe.shape[0] * E.shape[0] / (ekms * 1e-3), 1) if ekms > 0 else None,
           "call_s_with_host_copies": round(ete, 4)}
    esub = np.arange(0, Oe.shape[0], 100)
    te = time.perf_counter()
    Weo = O.winding_number_2d(P, E, Oe[esub])
    ecdt = time.perf_counter() - te
    ex2["max_abs_dw_vs_oracle"] = float(np.abs(We[esub] - Weo).max())
    if not no_cpu:
        ex2["cpu_baseline"] = {"value": round(esub.size / ecdt, 1), "unit": "queries/s", "cores": 1, "kind": "port",
                               "sample": f"{esub.size} queries, signed_angle summed over every edge (restatement)"}
    out["exact_2d"] = ex2
    if not no_cpu:
        # the reference has no parallel 2D wrapper: computeAngle called point by
        # point, as a caller would, on one core
        if O.ref_available():
            ref, kind = O.RefCurve(P, E), "reference"
        else:
            ref, kind = co, "port"
        csub = np.arange(0, n, 4)
        t = time.perf_counter()
        Ar = ref.angle(Qf[csub], 2.0) if kind == "reference" else ref.angle(Qf[csub], 2.0, 1)
        cdt = time.perf_counter() - t
        out["matches_reference_sample_max_abs_dw"] = float(np.abs(A[csub].astype(np.float64) - Ar).max() / (2 * np.pi))
        out["cpu_baseline"] = {"value": round(csub.size / cdt, 1), "unit": "queries/s", "cores": 1, "kind": kind,
                               "sample": f"{csub.size} queries (every 4th) of the same workload, computeAngle "
                                         "per point (serial)"}
    log(f"[bench] curve: {out}")
    return out


def run_serving(duration, beta_acc):
    """64 resident meshes (log-uniform 1k-500k triangles), Poisson arrivals at
    2000/s, batches log-uniform in [1k, 1M] uniform in the mesh's bbox; mixed
    precision.  Latency = scheduled arrival -> result in host memory.  Also
    steps the arrival rate up at accuracy_scale 2 and at the 1e-4 accuracy scale to
    find the largest rate the server sustains (throughput within 3% of offered,
    p99 under 50 ms)."""
    import numpy as np

    from libigl_amd import datagen as D
    from libigl_amd import gwn
    from libigl_amd.serving import Server, poisson_trace

    t0 = time.time()
    meshes = D.serving_meshes()
    srv = Server(slots=8, device_budget_bytes=64 << 30)
    ids = [srv.add_mesh(V, F, 2, gwn.TREE_SAH4) for V, F in meshes]
    srv.reserve(1_000_000)
    rng = np.random.default_rng(77)
    batches = [D.uniform_queries(V.min(0), V.max(0), 1_000_000, rng) for V, _ in meshes]
    ntris = sum(F.shape[0] for _, F in meshes)
    log(f"[bench] serving: 64 meshes, {ntris} tris resident, setup {time.time() - t0:.1f}s")
    res = {}

    def trace(beta, lam, dur):
        r = poisson_trace(srv, ids, batches, lam=lam, duration=dur, seed=9, accuracy_scale=beta,
                          precision=gwn.PREC_MIXED)
        log(f"[bench] serving beta {beta} lambda {lam}: {r}")
        return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}

    trace(2.0, 2000.0, 0.5)  # warm-up: first launches, page-in of the query batches
    for beta in (2.0, beta_acc):
        res[f"beta_{beta}"] = trace(beta, 2000.0, duration)
    stable = None
    for lam in (2000.0, 3000.0, 4000.0, 5000.0, 6000.0, 8000.0):
        r = trace(2.0, lam, min(duration, 1.0))
        if r["queries_per_s"] < 0.97 * r["offered_queries_per_s"] or r["p99_ms"] > 50.0:
            break
        stable = {"lambda": lam, "p99_ms": r["p99_ms"], "queries_per_s": r["queries_per_s"]}
    res["max_stable_lambda_beta_2"] = stable
    stable = None
    for lam in (250.0, 500.0, 750.0, 1000.0, 1500.0, 2000.0):
        r = trace(beta_acc, lam, min(duration, 1.0))
        if r["queries_per_s"] < 0.97 * r["offered_queries_per_s"] or r["p99_ms"] > 50.0:
            break
        stable = {"lambda": lam, "p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"], "queries_per_s": r["queries_per_s"]}
    res[f"max_stable_lambda_beta_{beta_acc}"] = stable
    cpu = None
    try:  # the reference's CPU path on a sample of the same requests, same host
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        threads = min(16, os.cpu_count() or 1)
        os.environ["IGL_NUM_THREADS"] = str(threads)
        if O.ref_available():
            r2 = np.random.default_rng(9)
            picks = r2.integers(0, len(meshes), size=8)
            nq_tot, t_tot = 0, 0.0
            for m in picks:
                Vm, Fm = meshes[m]
                ref = O.RefSoup(Vm, Fm)
                nqs = int(np.exp(r2.uniform(np.log(1e3), np.log(1e6))))
                Qm = batches[m][:nqs]
                t = time.perf_counter()
                ref.query(Qm, 2.0)
                t_tot += time.perf_counter() - t
                nq_tot += nqs
            cpu = {"value": round(nq_tot / t_tot, 1), "unit": "queries/s", "cores": threads, "kind": "reference",
                   "sample": f"8 requests of the trace's distribution ({nq_tot} queries), accuracy_scale 2, "
                             "UT_SolidAngle + igl::parallel_for, trees prebuilt"}
            log(f"[bench] serving CPU reference: {cpu}")
    except Exception as e:
        log(f"[bench] serving cpu baseline skipped: {e}")
    res["cpu_baseline"] = cpu
    resident, held, budget = srv.memory()
    srv.close()
    res.update({"meshes": 64, "resident_triangles": int(ntris), "resident_bytes": int(resident), "slots": 8,
                "slot_buffer_bytes": int(held), "lambda_per_s": 2000.0, "duration_s": duration,
                "precision": "mixed"})
    return res


if __name__ == "__main__":
    main()
