"""LINPACK benchmark line for the MI355X backend.

Headline workload (one GPU): mixed precision + FP64 iterative refinement,
n = 65536, block 256: BF16 trailing-update GEMM with FP32 accumulation, FP32
panel LU and strip solves, FP64 residual and correction on the GPU, on
A_ii = d * sum_{j != i} |A_ij| with d = 1.01, A_ij ~ U[-0.5,0.5) and
b = A x_true (x_true ~ U[-1,1), FP64), generated directly in HBM.

A step is one refined solve of the resident system (the matrix is restored
from an untouched device copy before each step, outside the timed events):
GEFA with the BF16 update, then GESL and classic refinement until the normwise
backward error ||b - A x|| / (||A|| ||x|| + ||b||) <= 1e-14 n.
value = effective GFLOP/s = (2/3 n^3 + 2 n^2) / t_solve averaged over the timed
steps -- the reference's metric (LINPACK/src/host/linpack_benchmark.hpp:
508-535) at FP64-level residual.  A step whose solve misses the backward-error
target fails the run: no value is reported for it.

The plain FP32 path (n = 32768, block 256, A_ii += n, b = A 1; the reference's
residn check) is the section "fp32_plain", with its own roofline.

Prints ONE JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MFMA_FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32 matrix peak (datasheet)
MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak (datasheet, no sparsity)

KERNEL_NAMES = {
    "fp32": "gemm_f32::update_kernel<128,128,16,32,64,2>",
    "bf16": "update_lpw_kernel<__bf16,256,256>",
}


def measured_traffic(name):
    """HBM bytes per trailing-update launch measured by rocprofv3 (FETCH_SIZE
    and WRITE_SIZE passes over the same workload, scaled by the counters'
    reading of a device copy of known size), or None when no measurement of
    this workload exists."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def _flops(m):
    return 2.0 / 3.0 * m ** 3 + 2.0 * m * m


def cpu_baseline(full_n, mode, d, seed=42, sizes=(1024, 2048, 4096), n_par=16384):
    """The reference CPU path (gefa_ref_nopvt + gesl_ref_nopvt restated in the
    oracle), one core, on bounded samples of the headline's generator: orders
    1024 (the reference's own CPU configuration), 2048 and 4096, with the
    t = c n^3 fit that extrapolates to the headline order.  Beside it the same
    loop nest with the independent columns of each step spread over the
    host's cores (a parallel restatement, not the reference) at n_par."""
    import numpy as np

    from oracle import oracle as O

    runs = []
    for n in sizes:
        A, b, _, xt = O.generate_sized(n, seed, mode, d)
        t0 = time.perf_counter()
        F = O.gefa_nopvt(A)
        t1 = time.perf_counter()
        x = O.gesl_nopvt(F, b)
        t2 = time.perf_counter()
        fwd = float(np.max(np.abs(x - xt)) / np.max(np.abs(xt)))
        runs.append({"n": n, "gefa_s": round(t1 - t0, 4), "gesl_s": round(t2 - t1, 5),
                     "gflops": round(_flops(n) / (t2 - t0) / 1e9, 3),
                     "forward_error_fp32": fwd})
    ns = np.array([r["n"] for r in runs], np.float64)
    ts = np.array([r["gefa_s"] + r["gesl_s"] for r in runs], np.float64)
    # t = c n^3 from the largest sample: the smaller ones run partly in cache
    # (n = 1024: 4 MB of matrix), so a fit over all of them would flatter the
    # extrapolation; their deviation from the n^3 law is reported beside it
    c = float(ts[-1] / ns[-1] ** 3)
    fit_dev = float(np.max(np.abs(c * ns ** 3 - ts) / ts))
    # the host share of the box: at most 16 threads
    threads = max(1, min(16, os.cpu_count() or 1))
    Ap, bp, _, xtp = O.generate_sized(n_par, seed, mode, d)
    t3 = time.perf_counter()
    Fp = O.gefa_nopvt_omp(Ap, threads)
    xp = O.gesl_nopvt(Fp, bp)
    t4 = time.perf_counter()
    fwdp = float(np.max(np.abs(xp - xtp)) / np.max(np.abs(xtp)))
    last = runs[-1]
    return {
        "value": last["gflops"],
        "unit": "GFLOP/s",
        "cores": 1,
        "kind": "port",
        "sample": f"gefa_ref_nopvt+gesl_ref_nopvt restated (oracle, g++ -O3 -march=x86-64-v3, "
                  f"1 thread) on the headline generator (mode {mode}, d={d}, seed {seed}) at "
                  f"n={last['n']}: gefa {last['gefa_s']}s gesl {last['gesl_s']}s, FP32 forward "
                  f"error {last['forward_error_fp32']:.2e} (no refinement on this path); cpu "
                  f"{_cpu_model()}, nproc {os.cpu_count()}",
        "runs_1core": runs,
        "t_eq_c_n3": {"c_s": c, "from_n": int(ns[-1]),
                      "max_rel_deviation_of_samples": round(fit_dev, 4)},
        "extrapolated_1core_s_at_n": {"n": full_n, "seconds": round(c * full_n ** 3, 1),
                                      "gflops": round(_flops(full_n) / (c * full_n ** 3) / 1e9, 3),
                                      "note": f"t = c n^3 scaled from n={int(ns[-1])}, not run"},
        "parallel_restatement": {"gflops": round(_flops(n_par) / (t4 - t3) / 1e9, 2),
                                 "threads": threads, "n": n_par,
                                 "seconds": round(t4 - t3, 2), "forward_error_fp32": fwdp,
                                 "note": "same loop nest, columns of each step over OpenMP "
                                         "threads; not the reference (single threaded). Not run "
                                         "at the headline order: at this rate it would take "
                                         f"{_flops(full_n) / (_flops(n_par) / (t4 - t3)):.0f} s"},
    }


def _roofline(kind, launches, gemm_s, gemm_flops, gemm_bytes, busy_s, peak, traffic_file):
    achieved = gemm_flops / gemm_s / 1e12 if gemm_s > 0 else 0.0
    device_achieved = gemm_flops / busy_s / 1e12 if busy_s > 0 else 0.0
    tr = measured_traffic(traffic_file)
    alg_bytes = gemm_bytes / max(1, launches)
    return {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": round(tr["hbm_bytes_per_launch"]) if tr else None,
            "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                            "copy-calibrated, profiles/" + traffic_file + ")",
            "traffic_over_algorithmic": round(tr["hbm_bytes_per_launch"] / alg_bytes, 3)
            if tr and alg_bytes > 0 else None,
            "algorithmic_bytes_per_launch": round(alg_bytes),
            "algorithmic_flops_per_launch": round(gemm_flops / max(1, launches)),
            "kernel": KERNEL_NAMES[kind],
            "launches": launches, "avg_launch_us": round(gemm_s / max(1, launches) * 1e6, 1),
            "kernel_s": round(gemm_s, 6),
            # achieved = algorithmic flops per launch / average launch duration
            # (HIP events on the launch streams).  The bulk update (stream s1),
            # the next panels' row and column-block updates (s2, and their
            # look-ahead on s3) and the diagonal blocks' updates (s0) run at
            # the same time, so each launch's duration is stretched by the
            # other streams' work and shared device time is counted once per
            # launch.  busy_union_* divide the same flops by the time during
            # which at least one launch runs (the union of their intervals):
            # the device's update rate.
            "achieved_definition": "algorithmic flops per launch / average launch duration",
            "busy_s": round(busy_s, 6),
            "busy_union_achieved": round(device_achieved, 3),
            "busy_union_frac": round(device_achieved / peak, 4),
            # The same kernel without the concurrent launches: rocprofv3's
            # counter runs (the FETCH_SIZE pass behind `traffic`) serialise the
            # dispatches, so their average duration is each launch's own time
            # in the same schedule.  From that committed profile, not this run.
            "serialised_achieved": round(tr["algorithmic_flops_per_launch"]
                                         / (tr["avg_kernel_ns_counters"] * 1e-9) / 1e12, 3)
            if tr else None,
            "serialised_frac": round(tr["algorithmic_flops_per_launch"]
                                     / (tr["avg_kernel_ns_counters"] * 1e-9) / 1e12 / peak, 4)
            if tr else None,
            "serialised_source": "profiles/" + traffic_file
                                 + " (algorithmic_flops_per_launch / avg_kernel_ns_counters)"}


def headline(ctx, n, B, d, steps, warmup, prec="bf16"):
    """The timed steps: refined mixed-precision solves of the resident system."""
    import numpy as np

    flops = _flops(n)
    tol = 1e-14 * n
    ctx.keep_original(True)
    tg = time.perf_counter()
    ctx.generate(n, 1, 42, d)
    print(f"[bench] generated n={n} d={d} in {time.perf_counter() - tg:.2f}s", file=sys.stderr,
          flush=True)
    for w in range(warmup):
        ctx.solve_ir(B, prec, max_iter=30, check=False)
        print(f"[bench] warmup {w} done", file=sys.stderr, flush=True)
    ctx.set_profiling(True)
    tot, lu, its, berr = [], [], [], []
    wall0 = time.perf_counter()
    for s in range(steps):
        r = ctx.solve_ir(B, prec, max_iter=30, check=False)
        tot.append(r["total_s"])
        lu.append(r["gefa_s"])
        its.append(r["iterations"])
        berr.append(r["backward_error"])
        print(f"[bench] step {s}: solve {r['total_s'] * 1e3:.2f} ms (gefa {r['gefa_s'] * 1e3:.2f}"
              f" ms, {r['iterations']} corrections, berr {r['backward_error']:.2e})",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - wall0
    launches, gemm_s, gemm_flops = ctx.kernel_stats()
    gemm_bytes = ctx.kernel_bytes()
    busy = ctx.kernel_busy()
    ctx.set_profiling(False)
    x = ctx.download_solution64()
    _, _, xt = ctx.download_vectors()
    fwd = float(np.max(np.abs(x - xt)) / np.max(np.abs(xt)))
    tot = np.array(tot)
    passed = bool(max(berr) <= tol and fwd <= 1e-10)
    return {
        "value": flops * len(tot) / float(np.sum(tot)) / 1e9,
        "tot": tot, "gefa": np.array(lu), "wall": wall,
        "validation": {"backward_error_max": max(berr), "tolerance": tol,
                       "forward_error_vs_x_true": fwd, "forward_error_bound": 1e-10,
                       "corrections": sorted(set(its)), "passed": passed},
        "roofline": _roofline(prec, launches, gemm_s, gemm_flops, gemm_bytes, busy,
                              MFMA_BF16_PEAK_TFLOPS, f"update_lp_bf16_n{n}_traffic.json"),
    }


def fp32_section(ctx, n, B, steps, warmup):
    """Plain FP32 GEFA + GESL, n = 32768, block 256, A_ii += n, b = A 1, with
    the reference's residn check and the HPL scaled residual."""
    import numpy as np

    flops = _flops(n)
    ctx.keep_original(True)
    ctx.generate(n, 0, 42)
    for _ in range(warmup):
        ctx.restore()
        ctx.gefa_dev(B)
        ctx.gesl_dev()
    ctx.set_profiling(True)
    t_lu, t_sl = [], []
    for _ in range(steps):
        ctx.restore()
        a = ctx.gefa_dev(B)
        b = ctx.gesl_dev()
        t_lu.append(a.gefa_s)
        t_sl.append(b.gesl_s)
    launches, gemm_s, gemm_flops = ctx.kernel_stats()
    gemm_bytes = ctx.kernel_bytes()
    busy = ctx.kernel_busy()
    ctx.set_profiling(False)
    rn, an, xn, bn = ctx.residual()
    _, x = ctx.download(matrix=False, solution=True)
    eps = float(np.finfo(np.float32).eps)
    residn = float(np.max(np.abs(x.astype(np.float64) - 1.0))) / (n * bn * eps)
    hpl = rn / (eps * an * xn * n)
    tot = np.array(t_lu) + np.array(t_sl)
    g = flops * len(tot) / float(np.sum(tot)) / 1e9
    print(f"[bench] fp32 n={n}: {g:.1f} GFLOP/s (gefa {np.mean(t_lu) * 1e3:.2f} ms, gesl "
          f"{np.mean(t_sl) * 1e3:.2f} ms)", file=sys.stderr, flush=True)
    return {"workload": f"plain FP32 LINPACK GEFA+GESL, n={n}, block {B}, 1 rhs, "
                        "U[-0.5,0.5) + n on the diagonal, b = A*1 (generator mode 0, seed 42)",
            "gflops": round(g, 2), "steps": steps, "ms_per_step": round(float(np.mean(tot)) * 1e3, 3),
            "gefa_ms_mean": round(float(np.mean(t_lu)) * 1e3, 3),
            "gesl_ms_mean": round(float(np.mean(t_sl)) * 1e3, 3),
            "gesl_gbps": round(4.0 * n * n / float(np.mean(t_sl)) / 1e9, 1),
            "validation": {"residn": residn, "hpl_scaled_residual": hpl,
                           "passed": bool(residn < 1 and hpl < 1)},
            "roofline": _roofline("fp32", launches, gemm_s, gemm_flops, gemm_bytes, busy,
                                  MFMA_FP32_PEAK_TFLOPS, f"gemm_f32_n{n}_traffic.json")}


def mixed_precision_section(ctx, n, B, fp64_oracle=True):
    """Mixed precision + FP64 refinement for d in {1.01, 1.1, 2.0}: the BF16
    update against the plain FP32 factorisation of the same matrix, each
    refined to 1e-14 n; forward errors against x_true and against the FP64 LU
    solution of the same system (the FP64 GEFA/GESL on the GPU)."""
    import numpy as np

    flops = _flops(n)
    out = {"workload": f"n={n}, block {B}, diagonal d*sum|A_ij|, b = A x_true (FP64)",
           "tolerance": f"backward error <= 1e-14*n = {1e-14 * n:.3e}", "runs": []}
    ctx.keep_original(True)
    for d in (1.01, 1.1, 2.0):
        row = {"d": d}
        x64 = None
        if fp64_oracle:
            # the same matrix in FP64 (the FP32 system widened), FP64 LU + GESL
            ctx.generate_f64(n, 1, 42, d)
            a = ctx.gefa_dev_f64()
            s = ctx.gesl_dev_f64()
            _, x64 = ctx.download_f64(matrix=False, solution=True)
            rn, an, xn, bn = ctx.residual_f64()
            row["fp64_lu"] = {"gflops": round(flops / (a.gefa_s + s.gesl_s) / 1e9, 1),
                              "backward_error": rn / (an * xn + bn)}
        else:
            ctx.generate(n, 1, 42, d)
        _, _, xt = ctx.download_vectors()
        for prec in ("fp32", "bf16"):
            ctx.solve_ir(B, prec, max_iter=30, check=False)  # warm-up of this shape
            r = ctx.solve_ir(B, prec, max_iter=30, check=False)
            x = ctx.download_solution64()
            row[prec] = {"gflops": round(flops / r["total_s"] / 1e9, 1),
                         "total_ms": round(r["total_s"] * 1e3, 2),
                         "gefa_ms": round(r["gefa_s"] * 1e3, 2), "iterations": r["iterations"],
                         "backward_error": r["backward_error"],
                         "forward_error_vs_x_true": float(np.max(np.abs(x - xt)) /
                                                          np.max(np.abs(xt))),
                         "converged": bool(r["converged"])}
            if x64 is not None:
                row[prec]["forward_error_vs_fp64_lu"] = float(np.max(np.abs(x - x64)) /
                                                              np.max(np.abs(x64)))
        row["speedup_vs_fp32"] = round(row["bf16"]["gflops"] / row["fp32"]["gflops"], 2)
        out["runs"].append(row)
        print(f"[bench] mixed d={d}: bf16 {row['bf16']['gflops']} GFLOP/s, fp32 "
              f"{row['fp32']['gflops']} GFLOP/s", file=sys.stderr, flush=True)
    return out


def pivoted_section(ctx, n, B, reps=3):
    """Partial pivoting (gefa_ref's path, --uniform): M = U[-0.5,0.5) with no
    diagonal dominance, b = M 1; FP32 GEFA+GESL, then FP64 refinement with the
    pivoted FP32 factor."""
    import numpy as np

    flops = _flops(n)
    ctx.keep_original(True)
    ctx.generate(n, 2, 42)
    ctx.set_pivoting(True)
    tot, lu = [], []
    for _ in range(reps + 1):
        ctx.restore()
        a = ctx.gefa_dev(B)
        b = ctx.gesl_dev()
        tot.append(a.gefa_s + b.gesl_s)
        lu.append(a.gefa_s)
    tot, lu = tot[1:], lu[1:]  # the first run warms up this shape
    rn, an, xn, _ = ctx.residual()
    eps = float(np.finfo(np.float32).eps)
    ir = ctx.solve_ir(B, "fp32", max_iter=10, check=False)
    ctx.set_pivoting(False)
    t = float(np.median(tot))
    print(f"[bench] pivoted n={n}: {flops / t / 1e9:.1f} GFLOP/s", file=sys.stderr, flush=True)
    return {"workload": f"partial pivoting, n={n}, block {B}, U[-0.5,0.5) (not diagonally "
                        f"dominant), b = A*1", "gflops": round(flops / t / 1e9, 1),
            "median_ms": round(t * 1e3, 3), "gefa_ms_median": round(float(np.median(lu)) * 1e3, 3),
            "hpl_scaled_residual": rn / (eps * an * xn * n),
            "refined": {"iterations": ir["iterations"], "backward_error": ir["backward_error"],
                        "converged": bool(ir["converged"]),
                        "total_ms": round(ir["total_s"] * 1e3, 3)}}


def fp64_section(ctx, n, reps=3):
    """The reference built with DATA_TYPE=double: the same generator's system
    in FP64, FP64 GEFA (FP64 MFMA update) + GESL, HPL residual in FP64."""
    import numpy as np

    flops = _flops(n)
    ctx.generate_f64(n, 0, 42)
    tot, lu = [], []
    for r in range(reps + 1):
        if r:
            ctx.restore_f64()
        a = ctx.gefa_dev_f64()
        b = ctx.gesl_dev_f64()
        tot.append(a.gefa_s + b.gesl_s)
        lu.append(a.gefa_s)
    tot, lu = tot[1:], lu[1:]
    rn, an, xn, _ = ctx.residual_f64()
    eps = float(np.finfo(np.float64).eps)
    t = float(np.median(tot))
    print(f"[bench] fp64 n={n}: {flops / t / 1e9:.1f} GFLOP/s", file=sys.stderr, flush=True)
    return {"workload": f"FP64 (DATA_TYPE=double build), n={n}, block 64, U[-0.5,0.5) + n on "
                        f"the diagonal, b = A*1", "gflops": round(flops / t / 1e9, 1),
            "peak_fp64_tflops": 78.6, "median_ms": round(t * 1e3, 3),
            "gefa_ms_median": round(float(np.median(lu)) * 1e3, 3),
            "hpl_scaled_residual_fp64": rn / (eps * an * xn * n)}


def batched_section(ctx, count=4096, seed=7):
    """4096 systems, n log-uniform in [32, 1024], FP16 trailing update + FP64
    refinement, one HIP-graph replay per solve; workgroup sizes autotuned."""
    import numpy as np

    from hpcc_fpga_amd.linpack import Batch

    u = np.array([_hash_uniform(seed, 1, i) for i in range(count)])
    orders = np.rint(np.exp(np.log(32) + u * (np.log(1024) - np.log(32)))).astype(np.int32)
    bt = Batch(ctx, orders)
    bt.generate(seed)
    threads = bt.autotune("fp16")
    times, st = [], None
    for _ in range(5):
        st = bt.solve("fp16")
        times.append(st["seconds"])
    bt.close()
    t = float(np.median(times))
    print(f"[bench] batched: {st['flops'] / t / 1e9:.1f} GFLOP/s", file=sys.stderr, flush=True)
    return {"workload": f"{count} systems, n log-uniform [32,1024] (seed {seed}), nrhs=1",
            "gflops": round(st["flops"] / t / 1e9, 1), "median_replay_ms": round(t * 1e3, 3),
            "max_iterations": st["max_iterations"],
            "max_backward_error_per_n": st["max_backward_error_per_n"],
            "failed": st["failed"], "autotuned_threads_per_bucket": [int(x) for x in threads]}


def serving_section(rate=None, count=600, seed=11, budget=200 * 10 ** 9):
    """Request loop: Poisson arrivals, n log-uniform in [64, 8192], nrhs in
    1..16, 20% of requests reusing a resident factor with new right-hand
    sides, FP8 E4M3 trailing update with per-block scales and FP64 refinement
    to 1e-14 n, HBM budget 200 GB.  Systems are generated on the device from
    per-request seeds; right-hand sides and solutions cross PCIe inside the
    measured latency.  With rate None the arrival rate is set to 70% of the
    throughput of a short closed-loop run."""
    import queue
    import threading

    import numpy as np

    from hpcc_fpga_amd import HplError, Server

    rng = np.random.default_rng(seed)
    try:
        srv = Server(slots=4, hbm_budget=budget, prec="fp8", max_n=8192, max_nrhs=16)
    except HplError as e:
        if e.status != 3:
            raise
        budget //= 2
        srv = Server(slots=4, hbm_budget=budget, prec="fp8", max_n=8192, max_nrhs=16)

    def draw():
        n = int(round(np.exp(rng.uniform(np.log(64), np.log(8192)))))
        return n, int(rng.integers(1, 17))

    pool, pool_lock = [], threading.Lock()  # (factor, n) of completed kept requests

    def run(arrivals):
        """Submit at the given offsets (s); returns (records, wall seconds)."""
        tickets = queue.Queue()
        recs = []

        def collector():
            while True:
                item = tickets.get()
                if item is None:
                    return
                t, kind, n, nrhs = item
                _, resp = srv.wait(t, check=False)
                recs.append((kind, n, nrhs, resp))
                if kind == "fresh" and resp["status"] == 0 and resp["factor"]:
                    with pool_lock:
                        pool.append((resp["factor"], n))
                        if len(pool) > 32:
                            f, _ = pool.pop(0)
                            srv.release(f)

        th = threading.Thread(target=collector)
        th.start()
        t0 = time.perf_counter()
        for k, at in enumerate(arrivals):
            dt = at - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            n, nrhs = draw()
            reuse = None
            if rng.random() < 0.2:
                with pool_lock:
                    if pool:
                        reuse = pool[int(rng.integers(len(pool)))]
            if reuse is not None:
                f, n = reuse
                b = rng.uniform(-1, 1, (n, nrhs))
                try:
                    t = srv.submit(b, factor=f)
                    tickets.put((t, "reuse", n, nrhs))
                    continue
                except HplError:
                    pass  # released meanwhile: serve it as a new system
                n, nrhs = draw()
            b = rng.uniform(-1, 1, (n, nrhs))
            t = srv.submit(b, seed=seed * 100000 + k, keep=True)
            tickets.put((t, "fresh", n, nrhs))
        tickets.put(None)
        th.join()
        return recs, time.perf_counter() - t0

    # closed-loop calibration: everything arrives at once
    cal, cal_wall = run([0.0] * 60)
    cap = len(cal) / cal_wall
    lam = rate if rate else 0.7 * cap
    arrivals = np.cumsum(rng.exponential(1.0 / lam, count))
    print(f"[bench] serving: capacity ~{cap:.0f} req/s, Poisson rate {lam:.0f} req/s, {count} "
          "requests", file=sys.stderr, flush=True)
    recs, wall = run(arrivals)
    st = srv.stats()
    srv.close()
    ok = [r for r in recs if r[3]["status"] == 0]
    lat = np.array([r[3]["latency_s"] for r in ok]) * 1e3
    flops = sum((2.0 / 3.0 * r[1] ** 3 + 2.0 * r[1] ** 2 * r[2]) if r[0] == "fresh"
                else 2.0 * r[1] ** 2 * r[2] for r in ok)
    out = {
        "workload": f"{count} requests, Poisson {lam:.1f}/s, n log-uniform [64,8192], nrhs 1..16, "
                    "20% factor reuse, FP8 E4M3 per-block-scaled update + FP64 refinement to "
                    "1e-14 n, device-generated systems (seeded), 4 slots",
        "hbm_budget_bytes": budget,
        "closed_loop_capacity_req_s": round(cap, 1),
        "arrival_rate_req_s": round(lam, 1),
        "solves_per_s": round(len(ok) / wall, 1),
        "effective_gflops": round(flops / wall / 1e9, 1),
        "p50_latency_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_latency_ms": round(float(np.percentile(lat, 99)), 3),
        "p50_queue_ms": round(float(np.percentile([r[3]["queue_s"] for r in ok], 50)) * 1e3, 3),
        "reuse_fraction": round(sum(r[0] == "reuse" for r in recs) / max(1, len(recs)), 3),
        "p50_latency_reuse_ms": round(float(np.percentile(
            [r[3]["latency_s"] * 1e3 for r in ok if r[0] == "reuse"] or [0.0], 50)), 3),
        "max_backward_error_per_n": max(r[3]["backward_error"] / r[1] for r in ok),
        "max_iterations": max(r[3]["iterations"] for r in ok),
        "failed": len(recs) - len(ok),
        "max_in_flight": st["max_in_flight"],
        "peak_factor_bytes": st["peak_factor_bytes"],
    }
    print(f"[bench] serving: p50 {out['p50_latency_ms']} ms p99 {out['p99_latency_ms']} ms, "
          f"{out['solves_per_s']} solves/s", file=sys.stderr, flush=True)
    return out


def _hash_uniform(seed, stream, k):
    """The counter-based uniform of the generators (splitmix64 finaliser)."""
    M = (1 << 64) - 1

    def mix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    key = mix(seed ^ (stream << 56))
    z = mix((key + (k + 1) * 0x9E3779B97F4A7C15) & M)
    return (z >> 40) / 16777216.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=65536, help="order of the headline solve")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--d", type=float, default=1.01, help="diagonal factor of the headline matrix")
    ap.add_argument("--fp32-n", type=int, default=32768, help="order of the plain FP32 section")
    ap.add_argument("--fp32-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the plain FP32 section")
    ap.add_argument("--only-fp32", action="store_true",
                    help="run the plain FP32 section alone and print it as the JSON line "
                         "(for profiling that section)")
    ap.add_argument("--only-batched", action="store_true",
                    help="run only the batched section (autotune, then five graph replays) and "
                         "print its JSON (for profiling)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the pivoted, FP64, mixed-precision sweep, batched and serving "
                         "sections")
    ap.add_argument("--serve-rate", type=float, default=None,
                    help="Poisson arrival rate (req/s); default 70%% of measured capacity")
    ap.add_argument("--serve-count", type=int, default=600)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU P x Q distribution is a design, not a measured path.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from hpcc_fpga_amd import Context

    n, B = args.n, args.block
    ctx = Context(0)
    dev = ctx.device_name
    print(f"[bench] device {dev}", file=sys.stderr, flush=True)
    if args.only_fp32:
        sec = fp32_section(ctx, args.fp32_n, B, args.fp32_steps, args.warmup)
        ctx.close()
        print(json.dumps({"fp32_plain": sec, "device": dev}), flush=True)
        return
    if args.only_batched:
        sec = batched_section(ctx)
        ctx.close()
        print(json.dumps({"batched": sec, "device": dev}), flush=True)
        return
    h = headline(ctx, n, B, args.d, args.steps, args.warmup)
    tot = h["tot"]

    extra = {}
    if not args.no_fp32:
        extra["fp32_plain"] = fp32_section(ctx, args.fp32_n, B, args.fp32_steps, 1)
    if not args.no_extra:
        extra["mixed_precision"] = mixed_precision_section(ctx, n, B)
        extra["pivoted"] = pivoted_section(ctx, args.fp32_n, 256)
        extra["fp64"] = fp64_section(ctx, args.fp32_n)
        extra["batched"] = batched_section(ctx)
    ctx.close()
    if not args.no_extra:
        extra["serving"] = serving_section(args.serve_rate, args.serve_count)

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(n, 1, args.d)

    passed = h["validation"]["passed"]
    line = {
        "metric": "effective GFLOP/s (2/3n^3+2n^2) at FP64-level residual; p50/p99 solve latency",
        # a run whose solves miss the FP64-level backward error reports no value
        "value": round(h["value"], 2) if passed else None,
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(np.mean(tot)) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": f"synthetic (counter-based U[-0.5,0.5), A_ii = {args.d} * sum_j |A_ij|, "
                "x_true ~ U[-1,1), b = A x_true in FP64, seed 42)",
        "config": {"workload": f"mixed-precision LINPACK + FP64 iterative refinement, n={n}, "
                               f"block {B}, d={args.d}, 1 rhs: BF16 trailing update with FP32 "
                               "accumulation, FP32 panel and solves, FP64 residual and "
                               "correction, refined to backward error <= 1e-14 n",
                   "n": n, "block": B, "d": args.d, "precision": "bf16+fp64-ir",
                   "generator": "sized mode 1"},
        "t_min_s": round(float(np.min(tot)), 6),
        "p50_solve_ms": round(float(np.percentile(tot, 50)) * 1e3, 3),
        "p99_solve_ms": round(float(np.percentile(tot, 99)) * 1e3, 3),
        "gefa_ms_mean": round(float(np.mean(h["gefa"])) * 1e3, 3),
        "wall_ms_per_step_incl_restore": round(h["wall"] / max(1, args.steps) * 1e3, 3),
        "validation": h["validation"],
        "roofline": h["roofline"],
        "cpu_baseline": cpu,
        "device": dev,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)
    if not passed:
        print("bench.py: the headline solves missed the FP64-level residual target; no value "
              "reported", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
