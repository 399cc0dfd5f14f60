#!/usr/bin/env python3
"""Benchmark of the accelerated nalgebra path on one MI355X.

Default run (`python bench.py`): one JSON line whose headline is the batched
Cholesky half of the metric, with the other workloads attached as records:

  headline       65,536 f64 SPD systems, n_j uniform integer in [4, 64],
                 A_j = G G^T + n_j I (G ~ N(0,1)), 4 right-hand sides ~ N(0,1);
                 Cholesky::new + solve_mut of every member (NROCM_EXACT, bit
                 for bit) in one hipGraph replay.  A step is one pass over a
                 fresh device-resident copy (prepared before the timed region).
  cholesky_fast  the same batch in NROCM_FAST (within 1e-12 of the oracle).
  lu_mixed_ir    n = 8192, kappa = 1e4: GPU FP64 LU (partial and windowed
                 pivoting) against FP32 / split-BF16 trailing updates + FP64
                 iterative refinement, all to backward error <= 1e-14.
  cholesky_adjoint, least_squares, lu_batched: the differentiated, least-
                 squares and general batched workloads.
  gemm           Matrix::gemm at 512^3 and 8192^3 (FP64 and FP32 MFMA).

    python bench.py [--workload all|chol|lu|adjoint|lsq|lu_batched|gemm]
                    [--steps K] [--warmup W] [--mode exact|fast] [--no-cpu]

`value` is matrices/s (the headline); every record carries its `roofline`
(algorithmic bytes or flops against the MI355X peak) and a `cpu_baseline`
timed with the oracle (a C restatement of nalgebra's loops) on this host.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "n=8192 LU-solve time, mixed+IR vs FP64 at equal error; batch Cholesky mats/s"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def host_info():
    """CPU model and core count of the machine the CPU baselines run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "cores_available": affinity}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--mode", choices=["exact", "fast"], default="exact")
    p.add_argument("--workload", choices=["all", "chol", "lu", "adjoint", "lsq", "lu_batched", "gemm"],
                   default="all")
    p.add_argument("--count", type=int, default=65536)
    p.add_argument("--seed", type=int, default=2)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-autotune", action="store_true", help="batched Cholesky: keep the default schedule")
    p.add_argument("--cpu-seconds", type=float, default=20.0)
    p.add_argument("--n", type=int, default=8192)
    p.add_argument("--reps-lu", type=int, default=3)
    p.add_argument("--warmup-lu", type=int, default=1)
    p.add_argument("--lu-variants", default="", help="comma-separated subset of the LU variants (default: all)")
    return p.parse_args()


def draw_sizes(lib, seed, count, lo=4, hi=64):
    u = np.zeros(count)
    lib.nrocm_gen_uniform(seed, count, 0.0, 1.0, u.ctypes.data_as(C.c_void_p))
    return np.minimum(lo + np.floor(u * (hi - lo + 1)), hi).astype(np.int32)


def gen_spd(lib, seed, ns, nrhs):
    n64 = ns.astype(np.int64)
    offA = np.concatenate([[0], np.cumsum(n64 * n64)[:-1]]).astype(np.int64)
    offB = np.concatenate([[0], np.cumsum(n64 * nrhs)[:-1]]).astype(np.int64)
    A = np.zeros(int((n64 * n64).sum()))
    B = np.zeros(int((n64 * nrhs).sum()))
    lib.nrocm_gen_spd_batch(seed, len(ns), ns.ctypes.data_as(C.POINTER(C.c_int)), nrhs,
                            A.ctypes.data_as(C.c_void_p), offA.ctypes.data_as(C.c_void_p),
                            B.ctypes.data_as(C.c_void_p), offB.ctypes.data_as(C.c_void_p))
    return A, offA, B, offB


def cpu_baseline_chol(ns, A, offA, B, offB, nrhs, budget_s):
    """Time the oracle's per-matrix Cholesky::new + solve_mut on the host: one
    thread, and one thread per core over member slices (nalgebra has no
    intra-factorisation threading; the oracle call releases the GIL).
    Each leg spends about budget_s / 2 seconds."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as orc

    def run(threads, budget):
        bounds = np.linspace(0, len(ns), threads + 1).astype(int)
        done, t_total, reps = 0, 0.0, 0
        with ThreadPoolExecutor(max_workers=threads) as ex:
            while t_total < budget and reps < 1000:
                Aw, Bw = A.copy(), B.copy()
                t0 = time.perf_counter()
                list(ex.map(lambda k: orc.potrf_potrs_batched(ns[bounds[k]:bounds[k + 1]], Aw,
                                                              offA[bounds[k]:bounds[k + 1]], Bw,
                                                              offB[bounds[k]:bounds[k + 1]], nrhs),
                            range(threads)))
                t_total += time.perf_counter() - t0
                done += len(ns)
                reps += 1
        return done / t_total, reps, t_total

    # at most 16 threads: the host share of one GPU on the measured machines
    cores = max(1, min(16, host_info()["cores_available"]))
    v1, r1, t1 = run(1, budget_s / 2)
    vn, rn, tn = run(cores, budget_s / 2)
    return {"value": vn, "unit": "matrices/s", "cores": cores, "kind": "port", **host_info(),
            "sample": f"the full {len(ns)}-matrix batch x{rn} (Cholesky::new + solve_mut per matrix, "
                      f"{cores} threads over member slices), {tn:.1f} s",
            "one_core": {"value": v1, "unit": "matrices/s", "cores": 1,
                         "sample": f"the full batch x{r1} on one thread, {t1:.1f} s"}}


def bench_chol(args):
    import torch
    import nalgebra_amd as nl
    from nalgebra_amd import _lib

    lib = _lib.lib()
    nrhs = 4
    ns = draw_sizes(lib, args.seed, args.count)
    A, offA, B, offB = gen_spd(lib, args.seed, ns, nrhs)
    n64 = ns.astype(np.int64)
    alg_bytes = int((8 * (n64 * (n64 + 1) + 2 * nrhs * n64)).sum())
    alg_flops = float((n64 ** 3 / 3.0 + 2.0 * nrhs * n64 ** 2).sum())
    mode = nl.EXACT if args.mode == "exact" else nl.FAST

    dev = torch.device("cuda:0")
    A0 = torch.from_numpy(A).to(dev)
    B0 = torch.from_numpy(B).to(dev)
    K = max(1, args.steps)
    copies = [(A0.clone(), B0.clone()) for _ in range(K)]
    bt = nl.CholeskyBatch(ns, nrhs, device=dev)
    ptrs = [bt.pointers(a, b) for a, b in copies]
    infos = [torch.empty(len(ns), dtype=torch.int32, device=dev) for _ in range(K)]
    # schedule per size bin: measured by the library's autotuner unless disabled
    t_tune = time.perf_counter()
    schedule = bt.schedule(mode) if args.no_autotune else bt.autotune(mode)
    t_tune = time.perf_counter() - t_tune

    # warm-up: capture one graph per copy, then restore the inputs
    for i in range(max(args.warmup, K)):
        j = i % K
        bt.solve_ptrs(ptrs[j][0], ptrs[j][1], infos[j], mode)
    torch.cuda.synchronize()
    for a, b in copies:
        a.copy_(A0)
        b.copy_(B0)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream(dev)
    ev0 = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    ev1 = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(K):
        ev0[i].record(stream)
        bt.solve_ptrs(ptrs[i][0], ptrs[i][1], infos[i], mode)
        ev1[i].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall_ms = (t1 - t0) * 1e3 / K
    step_ms = [ev0[i].elapsed_time(ev1[i]) for i in range(K)]
    kern_ms = float(np.mean(step_ms))

    bad = sum(int(x.count_nonzero().item()) for x in infos)
    if bad:
        raise SystemExit(f"{bad} members reported a non-positive pivot")
    # spot-check the last step against the oracle on a sample of members
    from oracle import oracle as orc
    Bg = copies[-1][1].cpu().numpy()
    sample = np.arange(0, len(ns), max(1, len(ns) // 512))
    Ar, Br = A.copy(), B.copy()
    sub_ns = ns[sample]
    orc.potrf_potrs_batched(sub_ns, Ar, offA[sample], Br, offB[sample], nrhs)
    max_rel = 0.0
    bit_equal = True
    for j in sample:
        sl = slice(offB[j], offB[j] + nrhs * ns[j])
        g, r = Bg[sl], Br[sl]
        max_rel = max(max_rel, float(np.linalg.norm(g - r) / np.linalg.norm(r)))
        bit_equal &= bool(np.array_equal(g.view(np.uint64), r.view(np.uint64)))

    # a headline figure is only published for results that keep the parity bar
    if args.mode == "exact" and not bit_equal:
        raise SystemExit("EXACT batched Cholesky is not bit-identical to the oracle on the sampled members")
    if not max_rel <= 1e-12:
        raise SystemExit(f"batched Cholesky relative error {max_rel:.3e} exceeds 1e-12")
    out = {
        "metric": METRIC,
        "value": len(ns) / (wall_ms * 1e-3),
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched_cholesky_solve", "count": int(len(ns)), "n_range": [4, 64],
                   "nrhs": nrhs, "mode": args.mode, "seed": args.seed, "graph": True,
                   "lanes_per_matrix": {str(k): v for k, v in schedule.items()},
                   "autotuned": not args.no_autotune, "autotune_s": round(t_tune, 3)},
        "roofline": {"bound": "hbm", "achieved": alg_bytes / (kern_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": alg_bytes / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "alg_bytes_per_step": alg_bytes},
        "gflops": alg_flops / (kern_ms * 1e-3) / 1e9,
        "gpu_event_ms_per_step": kern_ms,
        "parity": {"sampled_members": int(len(sample)), "max_rel_err": max_rel, "bit_identical": bit_equal},
    }
    # measured HBM bytes per step (rocprofv3 FETCH_SIZE + WRITE_SIZE passes of this
    # workload, calibrated against copies of known size), when recorded
    tr_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                           "batched_cholesky_65536_traffic.json")
    if os.path.exists(tr_path):
        with open(tr_path) as f:
            tr = json.load(f)
        if (tr.get("count"), tr.get("nrhs"), tr.get("seed"), tr.get("mode")) == (len(ns), nrhs, args.seed, args.mode):
            out["roofline"]["traffic"] = tr["traffic_bytes_per_step"]
            out["roofline"]["traffic_source"] = tr["summary"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_chol(ns, A, offA, B, offB, nrhs, args.cpu_seconds)
    return out


def haar_kappa_matrix(lib, n, seed, kappa=1e4):
    """A = U diag(sigma) V^T, U and V Haar-random (QR of a Gaussian with the
    sign fix of Mezzadri 2007, computed with torch on the GPU: input
    generation only), sigma log-spaced from 1 to 1/kappa."""
    import torch
    dev = torch.device("cuda:0")
    mats = []
    for s in (seed, seed + 1):
        g = np.zeros(n * n)
        lib.nrocm_gen_normal(s, n * n, g.ctypes.data_as(C.c_void_p))
        G = torch.from_numpy(g.reshape(n, n)).to(dev)
        q, r = torch.linalg.qr(G)
        mats.append(q * torch.sign(torch.diagonal(r)).unsqueeze(0))
    U, V = mats
    sig = torch.logspace(0, -np.log10(kappa), n, dtype=torch.float64, device=dev)
    A = (U * sig.unsqueeze(0)) @ V.T.contiguous()
    return A.T.contiguous().reshape(-1)   # column-major storage of A


def bench_lu(args):
    """Config 3: solve A x = b (n = 8192, kappa = 1e4) by FP64 LU on the GPU
    and by FP32 / split-BF16 trailing-update LU + FP64 iterative refinement,
    all to normwise backward error <= 1e-14.  value = mixed (FP32) time."""
    import torch
    import nalgebra_amd as nl
    from nalgebra_amd import _lib
    lib = _lib.lib()
    n = args.n
    dev = torch.device("cuda:0")
    A = haar_kappa_matrix(lib, n, args.seed)
    bh = np.zeros(n)
    lib.nrocm_gen_normal(args.seed + 7, n, bh.ctypes.data_as(C.c_void_p))
    b = torch.from_numpy(bh).to(dev)
    Am = nl.DMatrix(A, n, n)
    bm = nl.DMatrix(b, n, 1)
    res = {}
    flops = 2.0 / 3.0 * n ** 3
    W = nl.PREC_WINDOW
    precs = [("fp64", nl.PREC_FP64), ("fp32", nl.PREC_FP32), ("bf16x3", nl.PREC_BF16X3), ("bf16", nl.PREC_BF16),
             ("fp64_window", nl.PREC_FP64 | W), ("fp32_window", nl.PREC_FP32 | W),
             ("bf16x3_window", nl.PREC_BF16X3 | W), ("fp32_window_gmres", nl.PREC_FP32 | W | nl.PREC_GMRES),
             ("bf16_window_gmres", nl.PREC_BF16 | W | nl.PREC_GMRES)]
    if args.lu_variants:
        keep = set(args.lu_variants.split(",")) | {"fp64"}
        precs = [(k, v) for k, v in precs if k in keep]
    A2 = A.view(n, n).T  # A[i, j] as a row-major view of the column-major storage
    nrmA = float(A2.abs().sum(dim=1).max())
    xs = {}
    for name, p in precs:
        times, iters, berr, ok = [], 0, 0.0, False
        for rep in range(args.warmup_lu + args.reps_lu):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x, iters, berr, ok = nl.solve_mixed_ir(Am, bm, p, tol=1e-14, maxit=60 if p & nl.PREC_GMRES else 12)
            torch.cuda.synchronize()
            if rep >= args.warmup_lu:
                times.append((time.perf_counter() - t0) * 1e3)
        ms = float(np.median(times))
        xv = x.data.view(-1)
        xs[name] = xv
        # normwise backward error recomputed here in FP64
        eta = float((b - A2 @ xv).abs().max() / (nrmA * xv.abs().max() + b.abs().max()))
        res[name] = {"ms": ms, "iters": iters, "berr": berr, "berr_recomputed": eta, "converged": ok,
                     "fell_back_to_partial_pivoting": bool(nl.last_ir_fallback()),
                     "factor_tflops_equiv": flops / (ms * 1e-3) / 1e12}
    # forward error of every variant against the GPU FP64 LU (partial pivoting) solution
    ref = xs["fp64"]
    for name in res:
        res[name]["forward_err_vs_fp64"] = float((xs[name] - ref).abs().max() / ref.abs().max())
    # speedups against the GPU FP64 LU with partial pivoting, and for the
    # windowed variants also against FP64 with the same pivoting
    t64 = res["fp64"]["ms"]
    for name in res:
        res[name]["speedup_vs_fp64"] = t64 / res[name]["ms"]
        ref = res.get("fp64_window", {})
        if "_window" in name and ref.get("converged"):
            res[name]["speedup_vs_fp64_window"] = ref["ms"] / res[name]["ms"]
    best = min((v["ms"], k) for k, v in res.items() if v["converged"] and not k.startswith("fp64"))
    for name, v in res.items():
        if v["converged"] and not v["berr_recomputed"] <= 2e-14:
            raise SystemExit(f"LU variant {name}: recomputed backward error {v['berr_recomputed']:.3e} > 2e-14")
    out = {
        "metric": METRIC, "value": best[0], "unit": "ms", "n_gpus": 1, "steps": args.reps_lu,
        "warmup": args.warmup_lu, "ms_per_step": best[0], "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "lu_mixed_ir", "n": n, "kappa": 1e4, "tol": 1e-14, "best": best[1],
                   "seed": args.seed},
        "speedup_vs_fp64_partial_pivoting": res["fp64"]["ms"] / best[0],
        "speedup_vs_fp64_same_pivoting": (res["fp64_window"]["ms"] / best[0]
                                          if "_window" in best[1] and "fp64_window" in res else None),
        "results": res,
        "roofline": {"bound": "mfma", "achieved": res[best[1]]["factor_tflops_equiv"],
                     "peak": FP64_PEAK_TFLOPS if best[1].startswith("fp64") else FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s",
                     "frac": res[best[1]]["factor_tflops_equiv"]
                     / (FP64_PEAK_TFLOPS if best[1].startswith("fp64") else FP32_PEAK_TFLOPS),
                     "traffic": None,
                     "note": "fastest converged mixed-precision variant end to end (factor + refinement), "
                             "2/3 n^3 flops, against the FP32 MFMA peak (its trailing updates run there)"},
    }
    if not args.no_cpu:
        from oracle import oracle as orc
        # a bounded sample: the largest leading block whose oracle LU fits the CPU
        # budget (LU::new takes about 65 s at n = 8192 and 8 s at 4096 on one core),
        # scaled by (n / ncpu)^3
        ncpu = n if args.cpu_seconds >= 70 else min(n, 4096 if args.cpu_seconds >= 10 else 2048)
        Ah = np.ascontiguousarray(A.view(n, n).T[:ncpu, :ncpu].cpu().numpy())
        t0 = time.perf_counter()
        orc.getrf(Ah)
        tc = time.perf_counter() - t0
        est = tc * (n / ncpu) ** 3
        out["cpu_baseline"] = {"value": est * 1e3, "unit": "ms" if ncpu == n else "ms (extrapolated)", "cores": 1,
                               "kind": "port", **host_info(),
                               "sample": (f"oracle LU::new of the same n={n} matrix, {tc:.1f} s" if ncpu == n else
                                          f"oracle LU::new of the leading {ncpu}x{ncpu} block took {tc:.2f} s; "
                                          f"scaled by (n/{ncpu})^3")}
    return out


def bench_gemm(args):
    """Matrix::gemm (C = A B, column-major DMatrix operands, entries uniform
    [-1, 1]) through nrocm_dgemm / nrocm_sgemm: m = n = k = 512 (the size the
    single-matrix parity case uses) and 8192.  A step is one product of
    device-resident operands.  value = FP64 TFLOP/s at 8192^3; parity against
    the oracle (nalgebra's gemv loop order) on the whole 512^3 product and on
    sampled columns of the 8192^3 one, to 1e-12 (FP64) / 1e-5 (FP32) relative."""
    import torch
    import nalgebra_amd as nl
    from nalgebra_amd import _lib
    from oracle import oracle as orc
    lib = _lib.lib()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)

    def operands(n, seed, dtype):
        a = np.zeros(n * n); b = np.zeros(n * n)
        lib.nrocm_gen_uniform(seed, a.size, -1.0, 1.0, a.ctypes.data_as(C.c_void_p))
        lib.nrocm_gen_uniform(seed + 1, b.size, -1.0, 1.0, b.ctypes.data_as(C.c_void_p))
        A = nl.DMatrix(torch.from_numpy(a).to(dev, dtype), n, n)
        B = nl.DMatrix(torch.from_numpy(b).to(dev, dtype), n, n)
        return a, b, A, B

    def timed(n, dtype, steps, warmup, seed):
        a, b, A, B = operands(n, seed, dtype)
        Cm = nl.DMatrix.zeros(n, n, dtype, dev)
        for _ in range(warmup):
            Cm.gemm(1.0, A, B, 0.0)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            Cm.gemm(1.0, A, B, 0.0)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        return ms, a, b, Cm

    def rel(x, y):
        return float(np.linalg.norm(x - y) / np.linalg.norm(y))

    res = {}
    seed = args.seed + 300
    for name, n, dtype, peak, tol in (("f64_512", 512, torch.float64, FP64_PEAK_TFLOPS, 1e-12),
                                      ("f64_8192", 8192, torch.float64, FP64_PEAK_TFLOPS, 1e-12),
                                      ("f32_8192", 8192, torch.float32, FP32_PEAK_TFLOPS, 1e-5)):
        steps = max(1, args.steps) if n <= 512 else max(1, min(args.steps, 10))
        ms, a, b, Cm = timed(n, dtype, steps * (20 if n <= 512 else 1), max(args.warmup, 1), seed)
        tf = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
        npdt = np.float64 if dtype == torch.float64 else np.float32
        Ah = a.reshape(n, n).T.astype(npdt); Bh = b.reshape(n, n).T.astype(npdt)
        cols = np.arange(n) if n <= 512 else np.arange(0, n, n // 8)
        ref = orc.gemm(npdt(1.0), Ah, np.asfortranarray(Bh[:, cols]))
        got = Cm.numpy()[:, cols]
        err = rel(got.astype(np.float64), ref.astype(np.float64))
        if not err <= tol:
            raise SystemExit(f"gemm {name}: relative error {err:.3e} vs the oracle exceeds {tol:g}")
        res[name] = {"n": n, "ms": ms, "tflops": tf, "peak_tflops": peak, "frac": tf / peak,
                     "parity": {"columns_checked": int(len(cols)), "rel_err_vs_oracle": err, "tol": tol}}
    head = res["f64_8192"]
    out = {"metric": METRIC, "value": head["tflops"], "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": head["ms"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "gemm", "m": 8192, "n": 8192, "k": 8192, "layout": "column-major",
                      "seed": seed},
           "results": res,
           "roofline": {"bound": "mfma", "achieved": head["tflops"], "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": head["frac"], "traffic": None,
                        "note": "2 n^3 flops per product against the dense FP64 MFMA peak"}}
    if not args.no_cpu:
        a, b, _, _ = operands(512, seed, torch.float64)
        Ah = a.reshape(512, 512).T.copy(); Bh = b.reshape(512, 512).T.copy()
        t0 = time.perf_counter(); reps = 0
        while time.perf_counter() - t0 < min(args.cpu_seconds, 10.0) and reps < 200:
            orc.gemm(1.0, Ah, Bh)
            reps += 1
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": reps * 2.0 * 512 ** 3 / tc / 1e12, "unit": "TFLOP/s", "cores": 1,
                               "kind": "port", **host_info(),
                               "sample": f"{reps} oracle gemm calls at 512^3 (nalgebra's gemv loop order, one "
                                         f"thread), {tc:.1f} s"}
    return out


def bench_lu_batched(args):
    """Batched uneven general systems: 65,536 f64 matrices, n_j uniform in
    [4, 64], entries uniform[-1, 1], 4 right-hand sides; one step = LU::new +
    LU::solve_mut of every member (nrocm_dgetrf_getrs_batched, EXACT) on a fresh
    device-resident copy.  value = matrices/s."""
    import torch
    import nalgebra_amd as nl
    from nalgebra_amd import _lib
    from nalgebra_amd.batched import _pointer_array
    lib = _lib.lib()
    nrhs, count = 4, args.count
    ns = draw_sizes(lib, args.seed + 100, count)
    n64 = ns.astype(np.int64)
    offA = np.concatenate([[0], np.cumsum(n64 * n64)[:-1]]).astype(np.int64)
    offB = np.concatenate([[0], np.cumsum(n64 * nrhs)[:-1]]).astype(np.int64)
    offP = np.concatenate([[0], np.cumsum(n64)[:-1]]).astype(np.int64)
    A = np.zeros(int((n64 * n64).sum())); B = np.zeros(int((n64 * nrhs).sum()))
    lib.nrocm_gen_uniform(args.seed + 101, A.size, -1.0, 1.0, A.ctypes.data_as(C.c_void_p))
    lib.nrocm_gen_uniform(args.seed + 102, B.size, -1.0, 1.0, B.ctypes.data_as(C.c_void_p))
    dev = torch.device("cuda:0")
    A0 = torch.from_numpy(A).to(dev); B0 = torch.from_numpy(B).to(dev)
    K = max(1, args.steps)
    copies = [(A0.clone(), B0.clone()) for _ in range(K)]
    bt = nl.LUBatch(ns, nrhs, device=dev)
    P = torch.empty(int(n64.sum()), dtype=torch.int32, device=dev)
    info = torch.empty(count, dtype=torch.int32, device=dev)
    ptrs = [(_pointer_array(a, offA), _pointer_array(P, offP), _pointer_array(b, offB)) for a, b in copies]
    mode = nl.EXACT if args.mode == "exact" else nl.FAST
    for i in range(max(args.warmup, 1)):
        bt.solve_ptrs(*ptrs[i % K], info, mode)
    torch.cuda.synchronize()
    for a, b in copies:
        a.copy_(A0); b.copy_(B0)
    stream = torch.cuda.current_stream(dev)
    e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for i in range(K):
        bt.solve_ptrs(*ptrs[i], info, mode)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / K
    # parity on a sample of members against the oracle's LU::new + solve_mut
    from oracle import oracle as orc
    sample = np.arange(0, count, max(1, count // 512))
    Ar, Br = A.copy(), B.copy()
    pr, infr = orc.getrf_getrs_batched(ns[sample], Ar, offA[sample], Br, offB[sample],
                                       np.concatenate([[0], np.cumsum(n64[sample])[:-1]]), nrhs)
    Bg = copies[-1][1].cpu().numpy(); Pg = P.cpu().numpy()
    bit, piv_ok, po = True, True, 0
    for k, j in enumerate(sample):
        n = ns[j]
        sl = slice(offB[j], offB[j] + nrhs * n)
        bit &= bool(np.array_equal(Bg[sl].view(np.uint64), Br[sl].view(np.uint64)))
        piv_ok &= bool(np.array_equal(Pg[offP[j]:offP[j] + n], pr[po:po + n]))
        po += n
    if args.mode == "exact" and not (bit and piv_ok):
        raise SystemExit("EXACT batched LU is not bit-identical to the oracle on the sampled members")
    alg_bytes = int((8 * (2 * n64 * n64 + 2 * nrhs * n64) + 4 * n64).sum())
    out = {"metric": METRIC, "value": count / (ms * 1e-3), "unit": "matrices/s", "n_gpus": 1, "steps": K,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "batched_lu_solve", "count": count, "n_range": [4, 64], "nrhs": nrhs,
                      "mode": args.mode, "seed": args.seed + 100},
           "roofline": {"bound": "hbm", "achieved": alg_bytes / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": alg_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                        "alg_bytes_per_step": alg_bytes,
                        "note": "read A, write the packed L\\U, read B, write X, write the pivots"},
           "parity": {"sampled_members": int(len(sample)), "bit_identical": bit, "pivots_identical": piv_ok}}
    if not args.no_cpu:
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < args.cpu_seconds and reps < 100:
            Aw, Bw = A.copy(), B.copy()
            orc.getrf_getrs_batched(ns, Aw, offA, Bw, offB, offP, nrhs)
            reps += 1
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": reps * count / tc, "unit": "matrices/s", "cores": 1, "kind": "port",
                               **host_info(),
                               "sample": f"the full {count}-member batch x{reps} through the oracle's LU::new + "
                                         f"solve_mut (one thread), {tc:.1f} s"}
    return out


def bench_adjoint(args):
    """Config 4: 4,096 SPD 32x32 systems, b ~ N(0,1), loss = sum |x|^2; one step =
    forward (Cholesky::new + solve_mut) + hand-written adjoint (dA, db)."""
    import torch
    import nalgebra_amd as nl
    from nalgebra_amd import _lib
    lib = _lib.lib()
    count, n = 4096, 32
    ns = np.full(count, n, dtype=np.int32)
    A, offA, B, offB = gen_spd(lib, 4, ns, 1)
    dev = torch.device("cuda:0")
    K = max(1, args.steps)
    A0 = torch.from_numpy(A).to(dev)
    B0 = torch.from_numpy(B).to(dev)
    copies = [(A0.clone(), B0.clone()) for _ in range(K)]
    bt = nl.CholeskyBatch(ns, 1, device=dev)
    ptrs = [bt.pointers(a, b) for a, b in copies]
    infos = [torch.empty(count, dtype=torch.int32, device=dev) for _ in range(K)]
    mode = nl.EXACT if args.mode == "exact" else nl.FAST
    schedule = bt.autotune(mode)   # the forward's schedule, measured as in the batched Cholesky record
    dX = [torch.empty_like(B0) for _ in range(K)]
    dA = [torch.empty_like(A0) for _ in range(K)]
    dB = [torch.empty_like(B0) for _ in range(K)]
    from nalgebra_amd.batched import _pointer_array
    bptrs = [(ptrs[i][0], ptrs[i][1], _pointer_array(dX[i], bt.offB), _pointer_array(dA[i], bt.offA),
              _pointer_array(dB[i], bt.offB)) for i in range(K)]

    def step(i):
        bt.solve_ptrs(ptrs[i][0], ptrs[i][1], infos[i], mode)
        torch.mul(copies[i][1], 2.0, out=dX[i])          # dLoss/dx = 2x
        bt.backward_raw(*bptrs[i])

    for i in range(max(args.warmup, K)):
        step(i % K)
    torch.cuda.synchronize()
    for a, b in copies:
        a.copy_(A0)
        b.copy_(B0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(K):
        step(i)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / K
    # parity on a sample against the oracle's forward + analytic adjoint
    from oracle import oracle as orc
    Lg = copies[-1][0].cpu().numpy(); Xg = copies[-1][1].cpu().numpy()
    dAg = dA[-1].cpu().numpy(); dBg = dB[-1].cpu().numpy()
    bit = True
    for j in range(0, count, 64):
        Lr, _ = orc.potrf(A[offA[j]:offA[j] + n * n].reshape(n, n, order="F"))
        xr = orc.potrs(Lr, B[offB[j]:offB[j] + n])[:, 0]
        dAr, dBr = orc.potrs_adjoint(Lr, xr, 2 * xr)
        bit &= bool(np.array_equal(dAg[offA[j]:offA[j] + n * n], dAr.ravel(order="F")))
        bit &= bool(np.array_equal(dBg[offB[j]:offB[j] + n], dBr))
    alg_bytes = count * 8 * (n * (n + 1) + 2 * n + n * (n + 1) // 2 + 3 * n + n * n)
    out = {"metric": METRIC, "value": count / (ms * 1e-3), "unit": "matrices/s (forward+backward)", "n_gpus": 1,
           "steps": K, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "cholesky_solve_adjoint", "count": count, "n": n, "mode": args.mode,
                      "forward_lanes_per_matrix": {str(k): v for k, v in schedule.items()}},
           "roofline": {"bound": "hbm", "achieved": alg_bytes / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": alg_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                        "note": "the 54 MB working set is Infinity-Cache resident; launch/latency bound"},
           "parity": {"sampled_members": count // 64, "bit_identical_vs_oracle_adjoint": bit}}
    if not args.no_cpu:
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < args.cpu_seconds:
            for j in range(count):
                Lr, _ = orc.potrf(A[offA[j]:offA[j] + n * n].reshape(n, n, order="F"))
                xr = orc.potrs(Lr, B[offB[j]:offB[j] + n])[:, 0]
                orc.potrs_adjoint(Lr, xr, 2 * xr)
                reps += 1
                if time.perf_counter() - t0 > args.cpu_seconds:
                    break
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": reps / tc, "unit": "matrices/s (forward+backward)", "cores": 1,
                               "kind": "port", **host_info(), "sample": f"{reps} members through the oracle (ctypes per call), "
                                                         f"{tc:.1f} s"}
    return out


def bench_lsq(args):
    """Config 5: 2,048 tall f32 systems, m uniform in [256, 2048], n = 64, entries
    N(0,1), b = A x_true + 1e-3 N(0,1); blocked WY QR on FP32 MFMA + one FP64
    refinement step; value = systems/s; error vs FP64 LAPACK lstsq on a sample."""
    import torch
    import nalgebra_amd as nl
    from nalgebra_amd import _lib
    lib = _lib.lib()
    count, n = 2048, 64
    ms_ = draw_sizes(lib, 5, count, 256, 2048)
    m64 = ms_.astype(np.int64)
    tot = int(m64.sum())
    Ah = np.zeros(tot * n)
    lib.nrocm_gen_normal(5, tot * n, Ah.ctypes.data_as(C.c_void_p))
    xt = np.zeros(count * n)
    lib.nrocm_gen_normal(6, count * n, xt.ctypes.data_as(C.c_void_p))
    noise = np.zeros(tot)
    lib.nrocm_gen_normal(7, tot, noise.ctypes.data_as(C.c_void_p))
    offA = np.concatenate([[0], np.cumsum(m64 * n)[:-1]])
    offb = np.concatenate([[0], np.cumsum(m64)[:-1]])
    bh = np.zeros(tot)
    for j in range(count):
        Aj = Ah[offA[j]:offA[j] + ms_[j] * n].reshape(ms_[j], n, order="F")
        bh[offb[j]:offb[j] + ms_[j]] = Aj @ xt[j * n:(j + 1) * n] + 1e-3 * noise[offb[j]:offb[j] + ms_[j]]
    A32 = Ah.astype(np.float32); b32 = bh.astype(np.float32)
    dev = torch.device("cuda:0")
    Ad = torch.from_numpy(A32).to(dev); bd = torch.from_numpy(b32).to(dev)
    bt = nl.LeastSquaresBatch(ms_, n, device=dev)
    ptrs = bt.pointers(Ad, bd)
    x = torch.empty(count * n, dtype=torch.float64, device=dev)
    info = torch.empty(count, dtype=torch.int32, device=dev)
    for _ in range(max(1, args.warmup)):
        bt.solve_ptrs(ptrs[0], ptrs[1], x, info, 1)
    torch.cuda.synchronize()
    K = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(K):
        bt.solve_ptrs(ptrs[0], ptrs[1], x, info, 1)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / K
    xg = x.view(count, n).cpu().numpy()
    worst = 0.0
    for j in range(0, count, 64):
        Aj = A32[offA[j]:offA[j] + ms_[j] * n].reshape(ms_[j], n, order="F").astype(np.float64)
        xr = np.linalg.lstsq(Aj, b32[offb[j]:offb[j] + ms_[j]].astype(np.float64), rcond=None)[0]
        worst = max(worst, float(np.linalg.norm(xg[j] - xr) / np.linalg.norm(xr)))
    flops = float((2 * m64 * n * n - 2 * n ** 3 / 3).sum())
    out = {"metric": METRIC, "value": count / (ms * 1e-3), "unit": "systems/s", "n_gpus": 1, "steps": K,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64 refinement", "data": "synthetic",
           "config": {"workload": "batched_qr_least_squares", "count": count, "m_range": [256, 2048], "n": n,
                      "refine": 1},
           "roofline": {"bound": "mfma", "achieved": flops / (ms * 1e-3) / 1e12, "peak": 157.3, "unit": "TFLOP/s",
                        "frac": flops / (ms * 1e-3) / 1e12 / 157.3, "traffic": None},
           "error_vs_fp64_lstsq": {"sampled_systems": count // 64, "max_rel": worst}}
    if not args.no_cpu:
        from oracle import oracle as orc
        t0 = time.perf_counter()
        done = 0
        while time.perf_counter() - t0 < args.cpu_seconds:
            j = done % count
            orc.qr_ls(A32[offA[j]:offA[j] + ms_[j] * n].reshape(ms_[j], n, order="F"),
                      b32[offb[j]:offb[j] + ms_[j]])
            done += 1
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": done / tc, "unit": "systems/s", "cores": 1, "kind": "port", **host_info(),
                               "sample": f"{done} systems (the batch in order, repeated) through the oracle's f32 "
                                         f"QR::new + q_tr_mul + upper solve, {tc:.1f} s"}
    return out


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (batched members would "
              "shard across GPUs by member range without communication). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "all":
        # the default line: the batched-Cholesky record (headline value), with
        # the n = 8192 mixed-precision LU record and the differentiated and
        # least-squares batches attached, each with its own roofline and CPU
        # baseline (smaller CPU budgets so that the whole run stays short)
        out = bench_chol(args)
        sub = argparse.Namespace(**vars(args))
        sub.cpu_seconds = min(args.cpu_seconds, 10.0)
        sub.steps, sub.warmup = min(args.steps, 10), min(args.warmup, 3)
        sub.seed = 3
        sub.lu_variants = sub.lu_variants or "fp64,fp64_window,fp32_window,bf16x3_window,fp32"
        out["lu_mixed_ir"] = bench_lu(sub)
        sub.seed = args.seed
        out["cholesky_adjoint"] = bench_adjoint(sub)
        out["least_squares"] = bench_lsq(sub)
        out["lu_batched"] = bench_lu_batched(sub)
        out["gemm"] = bench_gemm(sub)
        fast = argparse.Namespace(**vars(args))
        fast.mode, fast.no_cpu = "fast", True
        out["cholesky_fast"] = bench_chol(fast)
        for k in ("lu_mixed_ir", "cholesky_adjoint", "least_squares", "lu_batched", "gemm", "cholesky_fast"):
            for f in ("metric", "n_gpus", "scaling", "vs_baseline", "data"):
                out[k].pop(f, None)
    elif args.workload == "chol":
        out = bench_chol(args)
    elif args.workload == "lu_batched":
        out = bench_lu_batched(args)
    elif args.workload == "gemm":
        out = bench_gemm(args)
    elif args.workload == "lu":
        out = bench_lu(args)
    elif args.workload == "adjoint":
        out = bench_adjoint(args)
    else:
        out = bench_lsq(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
