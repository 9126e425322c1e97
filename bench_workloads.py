"""The other named configurations, run through `python bench.py --workload NAME`.

    web     web-like 281,903^2 C = A*A, FP64, 4096-row blocks (literal Zipf 2.1 out-degree and
            the nnz-matched exponent 1.66), all three windows, the D2H and files windows again
            with the hybrid CPU co-execution mode on (CPU share reported), full-size oracle on the
            host
    rmat22  R-MAT scale 22 (a,b,c,d = .57,.19,.19,.05), C = A*A, FP64, 512-row blocks, 64 GB
            device budget: kernel window over the whole product (streamed in chunks), the whole
            product copied to pinned host memory (mode token d2hfull), D2H and file windows on
            the trailing 2^19 rows
    batch   512 independent C_i = A_i * B_i (FP32; rows log-uniform in [1e3, 2e5], mean
            nnz/row U[4, 64] with Zipf(1.8) row skew): one captured graph vs per-matrix plain runs
    fwdbwd  R-MAT scale 18 A (row-normalised adjacency + self loops) times an independent R-MAT
            B, FP32, through torch.autograd: forward vs forward + backward
    rocsparse  rocSPARSE SpGEMM (whole matrix and per 512-row block) beside DBD's kernel window
            on the web-like, Erdos-Renyi 2M and R-MAT 20 slice products, DBD's result checked
            against rocSPARSE's

Each prints one JSON line on stdout.
"""
import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(msg):
    print(f"[bench:{os.getpid()}] {msg}", file=sys.stderr, flush=True)


def _threads():
    t = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    return max(1, min(t, os.cpu_count() or 1))


def _timed(ctx, mode, steps, warmup):
    import torch
    for _ in range(warmup):
        ctx.run(mode)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    last = None
    for _ in range(steps):
        last = ctx.run(mode)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, last


def _windows(A, B, block_rows, dtype, steps, warmup, modes, device_gib=32, host_gib=48, io_dir=None):
    import dbd_amd
    out = {}
    ctx = dbd_amd.Context(block_rows=block_rows, dtype=dtype, device_gib=device_gib, host_gib=host_gib)
    ctx.set_operands(A, B)
    info = ctx.plan()
    P = info["products"]
    t, tm = _timed(ctx, "compute", steps, warmup)
    info = ctx.plan_info()
    out["compute"] = {"ms": round(t * 1e3, 3), "gflops": round(2 * P / t / 1e9, 3), "chunks": tm["chunks"]}
    if "d2h" in modes:
        t, tm = _timed(ctx, "d2h", max(1, min(steps, 3)), 1)
        out["d2h"] = {"ms": round(t * 1e3, 3), "gflops": round(2 * P / t / 1e9, 3), "d2h_bytes": tm["d2h_bytes"]}
    ctx.close()
    if "io" in modes:
        d = io_dir or tempfile.mkdtemp(prefix="dbd_io_", dir="/tmp")
        c2 = dbd_amd.Context(block_rows=block_rows, dtype=dtype, device_gib=device_gib, host_gib=host_gib,
                             out_dir=d, writer_threads=8, chunks_hint=16)
        c2.set_operands(A, B)
        c2.run("io")
        for f in ("Ptr.output", "Idx.output", "Val.output"):  # the timed run writes fresh files, as the CPU path does
            os.remove(os.path.join(d, f))
        t0 = time.perf_counter()
        tm = c2.run("io")
        t = time.perf_counter() - t0
        out["io"] = {"ms": round(t * 1e3, 3), "gflops": round(2 * P / t / 1e9, 3), "io_bytes": tm["io_bytes"],
                     "write_s": round(tm["write_s"], 3)}
        c2.close()
        shutil.rmtree(d, ignore_errors=True)
    return P, info, out


def _hybrid_windows(A, B, block_rows, dtype, steps, io=True, cpu_threads=0):
    """The D2H and files windows with the hybrid mode on: the blocks the reference's
    analyze_Block sends to the CPU (average products per row <= 64) run on host threads. Reports
    the CPU's share of blocks and products and its busy time beside the windows."""
    import dbd_amd
    out = {}
    ctx = dbd_amd.Context(block_rows=block_rows, dtype=dtype, device_gib=32, host_gib=48, hybrid=True,
                          cpu_threads=cpu_threads)
    ctx.set_operands(A, B)
    info = ctx.plan()
    P = info["products"]

    def share(tm):
        return {"cpu_blocks": tm["cpu_blocks"], "blocks": info["num_blocks"],
                "cpu_products_share": round(tm["cpu_products"] / max(P, 1), 4), "cpu_s": round(tm["cpu_s"], 4),
                "compute_gpu_s": round(tm["compute_gpu_s"], 4)}

    t, tm = _timed(ctx, "d2h", max(1, min(steps, 3)), 1)
    out["d2h"] = {"ms": round(t * 1e3, 3), "gflops": round(2 * P / t / 1e9, 3), **share(tm)}
    ctx.close()
    if io:
        d = tempfile.mkdtemp(prefix="dbd_hio_", dir="/tmp")
        c2 = dbd_amd.Context(block_rows=block_rows, dtype=dtype, device_gib=32, host_gib=48, out_dir=d,
                             writer_threads=8, chunks_hint=16, hybrid=True, cpu_threads=cpu_threads)
        c2.set_operands(A, B)
        c2.run("io")
        for f in ("Ptr.output", "Idx.output", "Val.output"):
            os.remove(os.path.join(d, f))
        t0 = time.perf_counter()
        tm = c2.run("io")
        t = time.perf_counter() - t0
        out["io"] = {"ms": round(t * 1e3, 3), "gflops": round(2 * P / t / 1e9, 3), **share(tm)}
        c2.close()
        shutil.rmtree(d, ignore_errors=True)
    return out


def _cpu_windows(A, B, threads, io=True):
    """The oracle as the CPU baseline: compute, + materialise the host CSR (already the
    oracle's output form), + write the same three files to the same kind of sink."""
    import oracle
    t0 = time.perf_counter()
    Cptr, Cidx, Cval = oracle.spgemm(A, B, threads=threads)
    t_c = time.perf_counter() - t0
    res = {"compute_s": round(t_c, 4)}
    if io:
        d = tempfile.mkdtemp(prefix="dbd_cpu_io_", dir="/tmp")
        t0 = time.perf_counter()
        (Cptr.astype(np.int32) if Cptr[-1] < 2 ** 31 else Cptr).tofile(os.path.join(d, "Ptr.output"))
        Cidx.tofile(os.path.join(d, "Idx.output"))
        Cval.tofile(os.path.join(d, "Val.output"))
        res["io_s"] = round(t_c + time.perf_counter() - t0, 4)
        shutil.rmtree(d, ignore_errors=True)
    return res, int(Cptr[-1])


def run_web(args):
    import torch

    import dbd_amd
    torch.cuda.set_device(0)
    modes = [m for m in args.modes.split(",") if m]
    rows = []
    for s in (2.1, 1.66):
        A = dbd_amd.gen_web(281903, zipf_s=s, seed=1)
        log(f"web s={s}: nnz {A.nnz}")
        P, info, win = _windows(A, None, 4096, np.float64, args.steps, args.warmup, modes)
        row = {"zipf_s": s, "nnz_A": A.nnz, "products": P, "nnz_C": info["nnz"], "windows": win,
               "blocks": {"criteria": info["criteria"], "gpu_only": info["blocks_gpu_only"],
                          "cpu_preferred": info["blocks_cpu_pref"], "any": info["blocks_any"]}}
        if "d2h" in modes:
            row["hybrid_windows"] = _hybrid_windows(A, None, 4096, np.float64, args.steps, io="io" in modes)
        if not args.no_cpu:
            th = _threads()
            cpu, _ = _cpu_windows(A, A, th, io="io" in modes)
            cpu1, _ = _cpu_windows(A, A, 1, io=False)
            row["cpu"] = {"threads": th, **cpu, "gflops": round(2 * P / cpu["compute_s"] / 1e9, 3),
                          "compute_s_1thread": cpu1["compute_s"]}
        rows.append(row)
        log(json.dumps(row))
    print(json.dumps({"metric": "SpGEMM GFLOPS (2*products/s) per window", "workload":
                      "web-like 281,903^2, C=A*A, FP64, 4096-row blocks", "unit": "GFLOPS", "runs": rows}))


def numeric_bytes(m, nnzA, P, nnzC, vsz=8):
    """Algorithmic HBM bytes of one numeric pass: A.ptr 8(m+1), A idx+val (4+v)nnzA, B.ptr pair
    8 nnzA, B idx+val gathers (4+v)P, C.ptr 8(m+1), C idx+val writes (4+v)nnzC."""
    return 8 * (m + 1) + (4 + vsz) * nnzA + 8 * nnzA + (4 + vsz) * P + 8 * (m + 1) + (4 + vsz) * nnzC


def rmat_kernel_window(scale=22, steps=1, warmup=1, A=None, table=None):
    """R-MAT C = A*A, FP64, 512-row blocks, 64 GB device budget: the kernel window over the whole
    product (chunks streamed within the budget), with the numeric pass's roofline (algorithmic
    bytes over the numeric pass's device time, summed over its chunks). `table`: a candidate
    category table id (default: the library's default table)."""
    import torch

    import dbd_amd
    t0 = time.perf_counter()
    if A is None:
        A = dbd_amd.gen_rmat(scale, 16, 0.57, 0.19, 0.19, seed=1)
    log(f"rmat{scale}: n={A.shape[0]} nnz={A.nnz} generated in {time.perf_counter() - t0:.1f}s")
    ctx = dbd_amd.Context(block_rows=512, dtype=np.float64, device_gib=64, host_gib=256, retain=False)
    if table is not None:
        ctx.set_table(table)
    ctx.set_operands(A)
    info = ctx.plan()
    P = info["products"]
    t, tm = _timed(ctx, "compute", steps, warmup)
    info = ctx.plan_info()
    nnzC = info["nnz"]
    nb = numeric_bytes(A.shape[0], A.nnz, P, nnzC)
    num_s = tm["numeric_s"]
    res = {"workload": f"R-MAT scale {scale}, edge factor 16, (.57,.19,.19,.05), C=A*A, FP64, 512-row blocks, "
                       f"device budget 64 GB",
           "n": A.shape[0], "nnz_A": A.nnz, "products": P, "nnz_C": nnzC,
           "kernel_window_ms": round(t * 1e3, 2), "gflops": round(2 * P / t / 1e9, 2), "chunks": tm["chunks"],
           "symbolic_pass_ms": round(tm["symbolic_s"] * 1e3, 2), "numeric_pass_ms": round(num_s * 1e3, 2),
           "numeric_roofline": {"bound": "hbm", "algorithmic_bytes": nb,
                                "achieved": round(nb / num_s / 1e9, 1), "peak": 8000.0, "unit": "GB/s",
                                "frac": round(nb / num_s / 1e9 / 8000.0, 4)},
           "num_rows_per_category": info["num_rows"]}
    if table is not None:
        res["table"] = table
    ctx.close()
    torch.cuda.empty_cache()
    log(f"rmat{scale}: kernel window {t * 1e3:.1f} ms ({2 * P / t / 1e9:.1f} GFLOPS), numeric pass "
        f"{num_s * 1e3:.1f} ms = {nb / num_s / 1e9:.0f} GB/s")
    return res, A


def rocsparse_run(args, timeout=1200):
    """Run the rocSPARSE harness program (dbd_amd/bin/rocsparse_spgemm) and return its JSON line.
    It times DBD's kernel window and rocSPARSE's SpGEMM (whole matrix, and per 512-row block as
    the reference's block harness runs it) on the same operands in its own process, and checks
    DBD's product against rocSPARSE's. A program of its own because a process that imports torch
    binds torch's bundled rocSPARSE and HIP runtime; its progress lines go to stderr as they come."""
    import subprocess
    exe = os.path.join(ROOT, "dbd_amd", "bin", "rocsparse_spgemm")
    r = subprocess.run([exe, *[str(x) for x in args]], stdout=subprocess.PIPE, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"rocsparse_spgemm {' '.join(map(str, args))} exited with {r.returncode}; "
                           f"stdout: {r.stdout[-2000:]!r}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def rocsparse_comparisons():
    """rocSPARSE beside DBD on the web-like 281,903^2 (FP64, DBD 4096-row blocks), the Erdos-Renyi
    2M^2 x 16 (FP64, this benchmark's workload) and an R-MAT scale 20 slice (rows 0, 16, 32, ... of
    A times A, FP64)."""
    import torch
    torch.cuda.empty_cache()  # the harness allocates in its own process
    res = {}
    log("rocSPARSE beside DBD: web-like 281,903^2")
    res["web"] = rocsparse_run(["--web", 281903, 1, "--dbd-block-rows", 4096])
    log("rocSPARSE beside DBD: Erdos-Renyi 2,000,000^2 x 16")
    res["er2m"] = rocsparse_run(["--er", 2_000_000, 2_000_000, 2_000_000, 16, 16, 1, 2, "--dbd-block-rows", 512])
    log("rocSPARSE beside DBD: R-MAT scale 20 slice")
    res["rmat20_slice"] = rocsparse_run(["--rmat-slice", 20, 16, "--dbd-block-rows", 512])
    return res


def run_rmat22(args):
    import torch

    import dbd_amd
    import oracle
    torch.cuda.set_device(0)
    sc = args.rmat_scale
    res, A = rmat_kernel_window(sc, max(1, args.steps), min(1, args.warmup), table=args.table)
    P = res["products"]
    modes = [m for m in args.modes.split(",") if m]
    if "d2hfull" in modes:
        # the whole product with C copied to pinned host memory (chunks through a 64 GB ring):
        # the end-to-end window whose sink is host memory, the CPU path's own output form
        c2 = dbd_amd.Context(block_rows=512, dtype=np.float64, device_gib=64, host_gib=64, retain=False)
        c2.set_operands(A)
        c2.plan()
        t2, tm2 = _timed(c2, "d2h", 1, 0)
        res["d2h_full"] = {"ms": round(t2 * 1e3, 2), "gflops": round(2 * P / t2 / 1e9, 2), "chunks": tm2["chunks"],
                           "d2h_bytes": tm2["d2h_bytes"],
                           "d2h_GBps": round(tm2["d2h_bytes"] / t2 / 1e9, 2)}
        log(json.dumps(res["d2h_full"]))
        c2.close()
    # D2H and file windows on a bounded contiguous row range (C of the full product is ~1 TB):
    # the last 2^19 rows, ~2.5e9 products (R-MAT's leading rows are its densest: the first
    # 2^18 rows alone hold 29% of all products)
    if "d2h" in modes or "io" in modes:
        r = min(A.shape[0], 1 << 19)
        r0 = A.shape[0] - r
        a0, a1 = int(A.ptr[r0]), int(A.ptr[-1])
        As = dbd_amd.CSR((A.ptr[r0:] - a0).copy(), A.idx[a0:a1].copy(), A.val[a0:a1].copy(), (r, A.shape[1]))
        Ps, infs, win = _windows(As, A, 512, np.float64, 1, 1, modes, device_gib=64, host_gib=64)
        res["trailing_rows"] = {"rows": [r0, A.shape[0]], "products": Ps, "nnz_C": infs["nnz"], "windows": win}
        log(json.dumps(res["trailing_rows"]))
        if not args.no_cpu:
            th = _threads()
            cpu, _ = _cpu_windows(As, A, th, io="io" in modes)
            res["trailing_rows"]["cpu"] = {"threads": th, **cpu, "gflops": round(2 * Ps / cpu["compute_s"] / 1e9, 3)}
    if not args.no_cpu:
        # CPU rate on a bounded sample of rows of the whole product
        Prow = oracle.row_products(A.ptr, A.idx, A.ptr)
        rng = np.random.default_rng(0)
        sel = np.sort(rng.choice(A.shape[0], 20000, replace=False)).astype(np.int32)
        t0 = time.perf_counter()
        oracle.spgemm_rows(sel, A, A, threads=_threads())
        dt = time.perf_counter() - t0
        res["cpu_sample"] = {"rows": int(sel.size), "products": int(Prow[sel].sum()), "s": round(dt, 3),
                             "gflops": round(2 * Prow[sel].sum() / dt / 1e9, 3), "threads": _threads()}
        # whole-product CPU time at the sample's rate (rows drawn uniformly at random)
        cpu_s = 2 * P / (res["cpu_sample"]["gflops"] * 1e9)
        res["cpu_sample"]["whole_product_s_est"] = round(cpu_s, 1)
        if "d2h_full" in res:
            res["end_to_end_d2h_vs_cpu"] = round(cpu_s / (res["d2h_full"]["ms"] * 1e-3), 2)
        res["kernel_vs_cpu"] = round(cpu_s / (res["kernel_window_ms"] * 1e-3), 2)
    print(json.dumps(res))


def _make_batch(n_mats, seed=2024, lo=1000, hi=200000):
    import dbd_amd
    rng = np.random.default_rng(seed)
    As, Bs = [], []
    for i in range(n_mats):
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        m = float(rng.uniform(4, 64))
        As.append(dbd_amd.gen_skewed(n, m, 1.8, seed=seed * 10007 + 2 * i, dtype=np.float32))
        Bs.append(dbd_amd.gen_skewed(n, m, 1.8, seed=seed * 10007 + 2 * i + 1, dtype=np.float32))
    return As, Bs


def batch_speedup(members=512, steps=1):
    """The batched form on `members` uneven FP32 products: one captured graph (category table
    autotuned on the concatenation) against the plain pipeline once per member with the same
    table. C is allocated at the exact nnz the batch's symbolic pass reports."""
    import torch

    import oracle
    from dbd_amd import batch
    t0 = time.perf_counter()
    As, Bs = _make_batch(members)
    A_cat, B_cat, lay = batch.concat(As, Bs)
    P = int(oracle.row_products(A_cat.ptr, A_cat.idx, B_cat.ptr).sum())
    log(f"batch of {len(As)}: rows {A_cat.shape[0]}, nnz A {A_cat.nnz}, products {P}, "
        f"built in {time.perf_counter() - t0:.1f}s")
    runner = batch.BatchRunner(A_cat, B_cat, block_rows=4096, tune=True)
    del A_cat, B_cat
    runner.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        runner.launch()
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / steps
    nnzC, table = runner.nnz, runner.table
    runner.close()
    del runner
    torch.cuda.empty_cache()
    log(f"batch: graph {t_batch * 1e3:.1f} ms, table {table}")
    def plain(tid):
        ctx, run = batch.plain_each(As, Bs, block_rows=4096, table=tid)
        run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ctx.close()
        torch.cuda.empty_cache()
        log(f"batch: plain per-matrix {dt * 1e3:.1f} ms (table {tid})")
        return dt

    t_plain = plain(table)
    res = {"workload": f"{len(As)} independent C_i = A_i*B_i, FP32, rows log-U[1e3,2e5], mean nnz/row U[4,64], "
                       f"Zipf(1.8)",
           "members": len(As), "products": P, "nnz_C": nnzC, "table": table,
           "batched_graph": {"ms": round(t_batch * 1e3, 2), "gflops": round(2 * P / t_batch / 1e9, 2)},
           "plain_per_matrix_same_table": {"ms": round(t_plain * 1e3, 2), "gflops": round(2 * P / t_plain / 1e9, 2)},
           "speedup": round(t_plain / t_batch, 3)}
    if table != 0:
        # the per-matrix runs with the default table as well: the tuned table is chosen for the
        # concatenation and need not suit the members run one by one
        t_def = plain(0)
        res["plain_per_matrix_default_table"] = {"ms": round(t_def * 1e3, 2), "gflops": round(2 * P / t_def / 1e9, 2)}
        res["speedup_vs_faster_plain"] = round(min(t_plain, t_def) / t_batch, 3)
    return res


def run_batch(args):
    import torch
    torch.cuda.set_device(0)
    res = batch_speedup(args.batch_members, max(1, args.steps))
    res["metric"] = "batched SpGEMM GFLOPS (2*products/s)"
    print(json.dumps(res))


def forward_backward(steps=3, warmup=1):
    """R-MAT scale 18 A (row-normalised adjacency + self loops) times an independent R-MAT B,
    FP32, through torch.autograd: forward alone vs forward + backward, loss sum(C * W)."""
    import torch

    import dbd_amd
    from dbd_amd.autograd import spgemm_values
    A = dbd_amd.gen_rmat(18, 16, seed=1, add_identity=True, row_normalize=True, dtype=np.float32)
    B = dbd_amd.gen_rmat(18, 16, seed=2, dtype=np.float32)
    Ad, Bd = dbd_amd.DeviceCSR.from_host(A), dbd_amd.DeviceCSR.from_host(B)
    av = Ad.val.clone().requires_grad_(True)
    bv = Bd.val.clone().requires_grad_(True)

    def fwd():
        return spgemm_values(Ad.ptr, Ad.idx, av, A.shape, Bd.ptr, Bd.idx, bv, B.shape, block_rows=512)

    Cv, Cp, Ci = fwd()
    W = torch.randn(Cv.shape, dtype=torch.float32, device="cuda")
    nnzC = int(Ci.numel())
    del Cv, Cp, Ci
    P = int(np.diff(B.ptr)[A.idx].sum())
    for _ in range(warmup):
        with torch.no_grad():
            fwd()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        with torch.no_grad():
            fwd()
    torch.cuda.synchronize()
    t_f = (time.perf_counter() - t0) / steps
    for _ in range(warmup):
        av.grad = bv.grad = None
        c, _, _ = fwd()
        (c * W).sum().backward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        av.grad = bv.grad = None
        c, _, _ = fwd()
        (c * W).sum().backward()
    torch.cuda.synchronize()
    t_fb = (time.perf_counter() - t0) / steps
    del W, c, av, bv, Ad, Bd
    torch.cuda.empty_cache()
    log(f"fwd {t_f * 1e3:.1f} ms, fwd+bwd {t_fb * 1e3:.1f} ms")
    return {"workload": "R-MAT scale 18 (row-normalised + self loops) x independent R-MAT scale 18, FP32, autograd",
            "nnz_A": A.nnz, "nnz_B": B.nnz, "products": P, "nnz_C": nnzC,
            "forward_ms": round(t_f * 1e3, 2), "fwd_bwd_ms": round(t_fb * 1e3, 2),
            "ratio": round(t_fb / t_f, 3), "forward_gflops": round(2 * P / t_f / 1e9, 2)}


def run_fwdbwd(args):
    import torch
    torch.cuda.set_device(0)
    res = forward_backward(max(1, args.steps), args.warmup)
    res["metric"] = "forward+backward / forward time"
    print(json.dumps(res))


def run_rocsparse(args):
    print(json.dumps(rocsparse_comparisons()), flush=True)
