"""Benchmark of the DBD SpGEMM hot path on one MI355X.

Workload (one "step" = one full C = A*B through the pipeline, planning included, inputs
already resident in HBM): A and B independent Erdos-Renyi 2,000,000 x 2,000,000 with exactly
16 distinct uniform columns per row, values N(0,1), FP64, 512-row blocks, device budget 32 GB.
P = 512,000,000 products; metric GFLOPS = 2 * P / t.

Windows (the reference's three binaries): kernel (`value`), kernel + D2H into pinned host
memory, kernel + D2H + file writes. The dominant kernel's roofline is measured with HIP
events on the stream it runs on; the CPU baseline is the oracle (C, OpenMP) over the whole
product, 1 warm-up + 5 runs (median), with a 1-thread figure beside it.

`extras` (skipped with --no-extras) carries the other configurations, measured in the same run:
R-MAT scale 22 C = A*A kernel window with the numeric pass's roofline, the R-MAT scale 18
forward+backward / forward ratio, and the 512-member batch against per-matrix plain runs with
the same tuned category table.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--modes compute,d2h,io] [--no-extras]
"""
import argparse
import json
import os
import shutil
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
N = 2_000_000
DEG = 16
BLOCK_ROWS = 512
DEVICE_GIB = 32


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def numeric_bytes(m, nnzA, P, nnzC, vsz=8):
    """Algorithmic HBM bytes of one numeric pass (int64 row pointers, int32 columns, v-byte
    values), independent of how the kernels lay out their workspace: A.ptr 8(m+1), A idx+val
    (4+v)nnzA, B.ptr pair 8 nnzA, B idx+val gathers (4+v)P, C.ptr 8(m+1), C idx+val writes
    (4+v)nnzC. Workspace the kernels add (row lists, per-nonzero B descriptors) is not counted."""
    return 8 * (m + 1) + (4 + vsz) * nnzA + 8 * nnzA + (4 + vsz) * P + 8 * (m + 1) + (4 + vsz) * nnzC


def symbolic_bytes(m, nnzA, P):
    """A.ptr 8(m+1), A idx 4 nnzA, B.ptr pair 8 nnzA, B idx gathers 4P, C row counts 8(m+1)."""
    return 8 * (m + 1) + 4 * nnzA + 8 * nnzA + 4 * P + 8 * (m + 1)


def cpu_info():
    """Model name, sockets and logical CPUs of this host (lscpu)."""
    import subprocess
    out = {}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except Exception:
        return out
    for line in txt.splitlines():
        k, _, v = line.partition(":")
        k, v = k.strip(), v.strip()
        if k == "Model name":
            out["cpu_model"] = v
        elif k == "Socket(s)":
            out["sockets"] = int(v) if v.isdigit() else v
        elif k == "CPU(s)":
            out["host_cpus"] = int(v) if v.isdigit() else v
    return out


def cpu_baseline(A, B, runs=5):
    """The oracle (row-wise Gustavson, C + OpenMP, FP64) on the whole product: 1 warm-up and
    `runs` timed runs on all usable host threads (median reported, min beside it), then one run
    on 1 thread over the first 1/16 of A's rows (the sequential convention of the reference's MKL
    baseline, mkl/Makefile:2)."""
    import statistics

    import dbd_amd
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(threads, os.cpu_count() or 1))
    P = int(oracle.row_products(A.ptr, A.idx, B.ptr).sum())
    oracle.spgemm(A, B, threads=threads)
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        oracle.spgemm(A, B, threads=threads)
        ts.append(time.perf_counter() - t0)
    med, mn = statistics.median(ts), min(ts)
    r1 = A.shape[0] // 16
    a1 = int(A.ptr[r1])
    As = dbd_amd.CSR(A.ptr[:r1 + 1].copy(), A.idx[:a1], A.val[:a1], (r1, A.shape[1]))
    P1 = int(oracle.row_products(As.ptr, As.idx, B.ptr).sum())
    t0 = time.perf_counter()
    oracle.spgemm(As, B, threads=1)
    t1 = time.perf_counter() - t0
    return {"value": round(2.0 * P / med / 1e9, 4), "unit": "GFLOPS", "cores": threads, "kind": "port",
            "sample": f"oracle C=A*B FP64 (row-wise Gustavson, OpenMP) over the whole product ({P} products), "
                      f"1 warm-up + {runs} runs on {threads} threads: median {med:.3f} s, min {mn:.3f} s",
            "median_s": round(med, 4), "min_s": round(mn, 4), "max_gflops": round(2.0 * P / mn / 1e9, 4),
            "one_thread": {"gflops": round(2.0 * P1 / t1 / 1e9, 4), "s": round(t1, 3),
                           "sample": f"rows [0, {r1}) of A with all of B ({P1} products)"},
            **cpu_info()}


def pinned_copy_gbps(nbytes=1 << 30, reps=3):
    import torch
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    dst = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
    dst.copy_(src, non_blocking=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src, non_blocking=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    del src, dst
    return nbytes / dt / 1e9


def load_traffic():
    path = os.path.join(ROOT, "profiles", "traffic_numeric_er2m.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        return d.get("bytes_per_launch")
    except Exception:
        return None


def profile_run():
    """Kernels for the counter passes: numeric pass x3 (one launch each), symbolic pass x1,
    then `y = x + 0` over 1 GiB of float32 x3 (reads 1 GiB, writes 1 GiB per launch)."""
    import numpy as np
    import torch

    import dbd_amd
    torch.cuda.set_device(0)
    A = dbd_amd.gen_er(N, N, DEG, seed=1)
    B = dbd_amd.gen_er(N, N, DEG, seed=2)
    ctx = dbd_amd.Context(block_rows=BLOCK_ROWS, dtype=np.float64, device_gib=DEVICE_GIB)
    ctx.set_operands(A, B)
    ctx.plan()
    Cptr = torch.empty(N + 1, dtype=torch.int64, device="cuda")
    nnz = ctx.symbolic(Cptr)
    Cidx = torch.empty(nnz, dtype=torch.int32, device="cuda")
    Cval = torch.empty(nnz, dtype=torch.float64, device="cuda")
    for _ in range(3):
        ctx.numeric(Cptr, Cidx, Cval)
    torch.cuda.synchronize()
    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.add(x, 0.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"profile_run": "done", "nnz_C": nnz, "copy_bytes_each_way": x.numel() * 4}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--modes", default="compute,d2h,io")
    ap.add_argument("--io-dir", default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--roofline-reps", type=int, default=5)
    ap.add_argument("--workload", default="er2m", choices=["er2m", "web", "rmat22", "batch", "fwdbwd", "rocsparse"],
                    help="er2m (default headline) or one of the other named configurations")
    ap.add_argument("--batch-members", type=int, default=512)
    ap.add_argument("--rmat-scale", type=int, default=22)
    ap.add_argument("--table", type=int, default=None,
                    help="rmat22 workload: run with this candidate category table (default: the default table)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the other configurations' lines (R-MAT 22 kernel window, forward+backward, "
                         "512-member batch) that the default run adds under \"extras\"")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3 counter passes: 3 numeric passes of the headline "
                         "workload, then 3 elementwise copies of 1 GiB (the known-size reference)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (each GPU box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload != "er2m":
        import bench_workloads
        getattr(bench_workloads, "run_" + args.workload)(args)
        return
    if args.profile:
        profile_run()
        return

    import numpy as np
    import torch

    import dbd_amd

    torch.cuda.set_device(0)
    modes = [m for m in args.modes.split(",") if m]
    log("generating A, B (ER 2M x 2M, 16/row, FP64)")
    t0 = time.perf_counter()
    A = dbd_amd.gen_er(N, N, DEG, seed=1)
    B = dbd_amd.gen_er(N, N, DEG, seed=2)
    log(f"generated in {time.perf_counter() - t0:.1f} s")

    ctx = dbd_amd.Context(block_rows=BLOCK_ROWS, dtype=np.float64, device_gib=DEVICE_GIB, host_gib=48,
                          retain=True, chunks_hint=8)
    ctx.set_operands(A, B)
    info = ctx.plan()
    P = info["products"]
    results = {}

    def timed(mode, steps, warmup):
        for _ in range(warmup):
            ctx.run(mode)
        torch.cuda.synchronize()
        ts = []
        t_all = time.perf_counter()
        last = None
        for _ in range(steps):
            ta = time.perf_counter()
            last = ctx.run(mode)
            ts.append(time.perf_counter() - ta)
        torch.cuda.synchronize()
        total = time.perf_counter() - t_all
        return total / steps, ts, last

    log(f"compute mode: {args.warmup} warmup + {args.steps} timed steps")
    t_c, ts_c, tm_c = timed("compute", args.steps, args.warmup)
    info = ctx.plan_info()
    nnzC = info["nnz"]
    results["compute"] = {"s": t_c, "gflops": 2 * P / t_c / 1e9, "steps_s": ts_c, "timing": tm_c}
    log(f"compute: {t_c * 1e3:.3f} ms/step, {2 * P / t_c / 1e9:.1f} GFLOPS, nnz(C)={nnzC}")

    if "d2h" in modes:
        st = max(1, min(args.steps, 3))
        log(f"d2h mode: 1 warmup + {st} steps")
        t_d, ts_d, tm_d = timed("d2h", st, 1)
        results["d2h"] = {"s": t_d, "gflops": 2 * P / t_d / 1e9, "timing": tm_d,
                          "d2h_bytes": tm_d["d2h_bytes"],
                          "effective_GBps": tm_d["d2h_bytes"] / max(t_d - t_c, 1e-9) / 1e9}
        log(f"d2h: {t_d * 1e3:.1f} ms/step, {2 * P / t_d / 1e9:.2f} GFLOPS")
    if "io" in modes:
        d = args.io_dir or tempfile.mkdtemp(prefix="dbd_io_", dir="/tmp")
        os.makedirs(d, exist_ok=True)
        ctx_io = dbd_amd.Context(block_rows=BLOCK_ROWS, dtype=np.float64, device_gib=DEVICE_GIB, host_gib=48,
                                 out_dir=d, writer_threads=8, chunks_hint=16)
        ctx_io.set_operands(A, B)
        log(f"io mode into {d}: 1 warmup + 1 step")
        ctx_io.run("io")
        for f in ("Ptr.output", "Idx.output", "Val.output"):  # the timed run writes fresh files, as the CPU path does
            os.remove(os.path.join(d, f))
        ta = time.perf_counter()
        tm_io = ctx_io.run("io")
        t_io = time.perf_counter() - ta
        fs = "?"
        try:
            import subprocess
            fs = subprocess.run(["df", "-T", d], capture_output=True, text=True).stdout.splitlines()[-1]
        except Exception:
            pass
        results["io"] = {"s": t_io, "gflops": 2 * P / t_io / 1e9, "timing": tm_io, "sink": d, "df": fs}
        log(f"io: {t_io * 1e3:.1f} ms, {2 * P / t_io / 1e9:.2f} GFLOPS, sink {fs}")
        ctx_io.close()
        shutil.rmtree(d, ignore_errors=True)

    # ---- roofline of the dominant kernel (numeric pass) with HIP events on its stream ----
    log("roofline: numeric pass timed with events on the library's stream")
    s = torch.cuda.Stream()
    ctx.set_stream(s)
    with torch.cuda.stream(s):
        Cptr = torch.empty(N + 1, dtype=torch.int64, device="cuda")
        nnz2 = ctx.symbolic(Cptr)
        Cidx = torch.empty(nnz2, dtype=torch.int32, device="cuda")
        Cval = torch.empty(nnz2, dtype=torch.float64, device="cuda")
        pinfo = ctx.plan_info()
        ncat = sum(1 for x in pinfo["num_rows"][1:] if x > 0)
        ctx.numeric(Cptr, Cidx, Cval)  # warm
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(args.roofline_reps):
            ctx.numeric(Cptr, Cidx, Cval)
        e1.record(s)
        e1.synchronize()
        num_ms = e0.elapsed_time(e1) / args.roofline_reps
        # symbolic pass (includes its binning and scans) for reference
        e0.record(s)
        for _ in range(args.roofline_reps):
            ctx.symbolic(Cptr)
        e1.record(s)
        e1.synchronize()
        sym_ms = e0.elapsed_time(e1) / args.roofline_reps
    torch.cuda.synchronize()
    nnzA = A.nnz
    nb_bytes = numeric_bytes(N, nnzA, P, nnz2)
    achieved = nb_bytes / (num_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": load_traffic(),
            "kernel": "%s<double,512,64,16> (numeric pass, %d category launch(es))"
                      % ("k_numeric_wave_fast" if 10 * nnz2 >= 9 * P else "k_numeric_wave", ncat),
            "kernel_ms": round(num_ms, 4), "algorithmic_bytes": nb_bytes,
            "symbolic_pass_ms": round(sym_ms, 4),
            "symbolic_achieved_GBps": round(symbolic_bytes(N, nnzA, P) / (sym_ms * 1e-3) / 1e9, 1)}
    log(f"numeric {num_ms:.3f} ms -> {achieved:.0f} GB/s ({achieved / HBM_PEAK_GBPS:.1%} of 8 TB/s); "
        f"symbolic pass {sym_ms:.3f} ms")
    del Cidx, Cval, Cptr
    ctx.close()

    extra = {}
    if not args.no_extras:
        import bench_workloads as bw
        ex = {}
        log("extras: R-MAT scale 22 kernel window (1 warm-up + 1 run)")
        ex["rmat22"], _ = bw.rmat_kernel_window(22, steps=1, warmup=1)
        log("extras: forward + backward, R-MAT scale 18, FP32")
        ex["fwd_bwd"] = bw.forward_backward(steps=3, warmup=1)
        log("extras: 512-member batch vs per-matrix plain runs (same tuned table)")
        ex["batch"] = bw.batch_speedup(512, steps=1)
        extra["extras"] = ex
    if "d2h" in modes:
        try:
            extra["pinned_copy_GBps"] = round(pinned_copy_gbps(), 2)
        except Exception as e:  # pragma: no cover
            extra["pinned_copy_GBps"] = f"error: {e}"
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle on host cores, bounded sample)")
        cpu = cpu_baseline(A, B)
        log(f"cpu: {cpu['value']} GFLOPS on {cpu['cores']} threads")

    out = {
        "metric": "SpGEMM GFLOPS (2*products/s), kernel window",
        "value": round(results["compute"]["gflops"], 3),
        "unit": "GFLOPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(results["compute"]["s"] * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, same shape as the named configuration)",
        "config": {"workload": "Erdos-Renyi A,B 2,000,000 x 2,000,000, 16 nnz/row, C=A*B, FP64, 512-row blocks, "
                               "device budget 32 GB",
                   "products": P, "nnz_C": nnzC, "block_rows": BLOCK_ROWS, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "windows": {k: {"ms": round(v["s"] * 1e3, 3), "gflops": round(v["gflops"], 3)} for k, v in results.items()},
    }
    if "d2h" in results:
        out["windows"]["d2h"]["d2h_bytes"] = results["d2h"]["d2h_bytes"]
        out["windows"]["d2h"]["effective_GBps"] = round(results["d2h"]["effective_GBps"], 2)
    if "io" in results:
        out["windows"]["io"]["sink"] = results["io"]["df"]
        out["windows"]["io"]["write_s"] = round(results["io"]["timing"]["write_s"], 3)
        out["windows"]["io"]["io_bytes"] = results["io"]["timing"]["io_bytes"]
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
