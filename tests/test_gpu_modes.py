"""The reference's three modes (bin/compute, bin/d2h, bin/SSpMM), budgets, chunking, the CLI
drop-ins and the device two-phase API, all against the oracle."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

import dbd_amd
from helpers import compare, gpu_spgemm, reference

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def test_d2h_mode_matches_oracle(gpu_ok):
    A = dbd_amd.gen_web(60000, zipf_s=1.66, seed=3)
    ref = reference(A)
    C, t, info = gpu_spgemm(A, mode="d2h", block_rows=512, chunks_hint=7)
    compare(C, ref, np.float64, "d2h")
    assert t["chunks"] >= 7 and t["d2h_bytes"] >= 12 * C.nnz


@pytest.mark.parametrize("dtype,ptr64", [(np.float64, False), (np.float32, True)])
def test_io_mode_files(gpu_ok, tmp_path, dtype, ptr64):
    A = dbd_amd.gen_rmat(13, 16, seed=4, dtype=dtype)
    ref = reference(A)
    C, t, info = gpu_spgemm(A, mode="io", block_rows=512, out_dir=str(tmp_path), ptr64=ptr64, chunks_hint=5,
                            writer_threads=3)
    p = np.fromfile(tmp_path / "Ptr.output", dtype=np.int64 if ptr64 else np.int32).astype(np.int64)
    i = np.fromfile(tmp_path / "Idx.output", dtype=np.int32)
    v = np.fromfile(tmp_path / "Val.output", dtype=dtype)
    compare(dbd_amd.CSR(p, i, v, A.shape), ref, dtype, "io files")
    assert t["io_bytes"] == (8 if ptr64 else 4) * (A.shape[0] + 1) + (4 + np.dtype(dtype).itemsize) * i.size
    assert t["chunks"] >= 5


def test_small_device_budget_forces_chunks(gpu_ok):
    """Output larger than the device budget left for it: compute mode streams chunks
    (results not retained, like the reference's kernel mode), D2H mode still returns all of C."""
    A = dbd_amd.gen_er(200000, 200000, 16, seed=5)
    B = dbd_amd.gen_er(200000, 200000, 16, seed=6)
    ref = reference(A, B)
    # inputs ~2 x 39 MB + workspace ~10 MB; C is ~614 MB -> a 300 MB budget forces chunks
    C, t, info = gpu_spgemm(A, B, mode="d2h", block_rows=512, device_gib=0.3)
    compare(C, ref, np.float64, "budget d2h")
    assert t["chunks"] > 2
    with dbd_amd.Context(block_rows=512, device_gib=0.3) as ctx:
        ctx.set_operands(A, B)
        t = ctx.run("compute")
        assert t["chunks"] > 1
        with pytest.raises(dbd_amd.DBDError) as e:
            ctx.result()
        assert e.value.status == 7


def test_budget_errors(gpu_ok):
    A = dbd_amd.gen_er(20000, 20000, 16, seed=7)
    with dbd_amd.Context(block_rows=512, device_gib=0.001) as ctx:  # 1 MiB: below the inputs
        ctx.set_operands(A)
        with pytest.raises(dbd_amd.DBDError) as e:
            ctx.run("compute")
        assert e.value.status == 4 and "Not Enough Memory" in str(e.value)
    with dbd_amd.Context(block_rows=512, host_gib=0.001) as ctx:  # host ring cannot hold 2 chunks
        ctx.set_operands(A)
        with pytest.raises(dbd_amd.DBDError) as e:
            ctx.run("io")
        assert e.value.status == 4


def test_io_error_is_reported(gpu_ok, tmp_path):
    A = dbd_amd.gen_er(5000, 5000, 8, seed=8)
    with dbd_amd.Context(block_rows=512, out_dir=str(tmp_path / "missing" / "dir")) as ctx:
        ctx.set_operands(A)
        with pytest.raises(dbd_amd.DBDError) as e:
            ctx.run("io")
        assert e.value.status == 5


def test_invalid_operands_rejected(gpu_ok):
    A = dbd_amd.gen_er(100, 100, 4, seed=9)
    bad = dbd_amd.CSR(A.ptr, A.idx[::-1].copy(), A.val, A.shape)  # unsorted rows
    with dbd_amd.Context() as ctx:
        with pytest.raises(dbd_amd.DBDError) as e:
            ctx.set_operands(bad)
        assert e.value.status == 1
        B = dbd_amd.gen_er(50, 100, 4, seed=9)
        with pytest.raises(dbd_amd.DBDError):
            ctx.set_operands(A, B)  # 100 x 100 times 50 x 100
        with pytest.raises(dbd_amd.DBDError):
            ctx.set_operands(A.astype(np.float32))  # context is FP64


@pytest.mark.parametrize("exe,mode", [("compute", 0), ("d2h", 1), ("SSpMM", 2), ("debug", 2)])
def test_cli_drop_in(gpu_ok, tmp_path, exe, mode):
    """The reference argv forms and report labels; SSpMM writes the reference file formats."""
    src = tmp_path / "lap.mtx"
    shutil.copy(os.path.join(GOLD, "laplacian_5pt_10x10.mtx"), src)
    binp = os.path.join(ROOT, "dbd_amd", "bin", exe)
    r = subprocess.run([binp, "--block-rows", "256", "4", "2", str(src)], capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    for label in ("Preprocessing", "Compute GPU", "Write File", "Kernel Execution", "Total Execution"):
        assert label in r.stdout
    assert (tmp_path / "lap.mtx.csr").exists()  # the reference's cache is written
    if exe == "debug":
        assert "- DEBUG -" in r.stdout and "Reduced to" in r.stdout
        # the reference's DEBUG lines (Z/distributeWorkload.cpp:42-43 before the run, Z/main.cpp:113 after)
        assert r.stdout.index("Top 10%") < r.stdout.index("Preprocessing")
        nb = int(re.search(r"Divide Matrix A into (\d+) Blocks", r.stdout).group(1))
        m = re.search(r"Merged (-?\d+) Blocks while Writing \(([0-9.]+)\)", r.stdout)
        assert m, r.stdout
        merged, ratio = int(m.group(1)), float(m.group(2))
        assert 0 <= merged < nb and abs(ratio - merged / nb) < 1e-6
    if mode == 2:
        A = dbd_amd.read_matrix(str(src), dtype=np.float32)
        ref = reference(A)
        p = np.fromfile(tmp_path / "Ptr.output", dtype=np.int32).astype(np.int64)
        i = np.fromfile(tmp_path / "Idx.output", dtype=np.int32)
        v = np.fromfile(tmp_path / "Val.output", dtype=np.float32)
        compare(dbd_amd.CSR(p, i, v, A.shape), ref, np.float32, exe)
    else:
        assert not (tmp_path / "Idx.output").exists()


def test_cli_a_times_b_fp64(gpu_ok, tmp_path):
    A = dbd_amd.gen_er(3000, 2000, 6, seed=10)
    B = dbd_amd.gen_er(2000, 2500, 5, seed=11)
    for name, M in (("a.mtx", A), ("b.mtx", B)):
        with open(tmp_path / name, "w") as f:
            f.write("%%MatrixMarket matrix coordinate real general\n")
            # sizes one less than the dims: the reference keeps indices as given (dims = header + 1)
            f.write(f"{M.shape[0] - 1} {M.shape[1] - 1} {M.nnz}\n")
            rows = np.repeat(np.arange(M.shape[0]), np.diff(M.ptr))
            for r_, c_, v_ in zip(rows, M.idx, M.val):
                f.write(f"{r_} {c_} {float(v_):.17g}\n")
    binp = os.path.join(ROOT, "dbd_amd", "bin", "SSpMM")
    r = subprocess.run([binp, "--fp64", "--no-cache", "--block-rows", "512", str(tmp_path / "a.mtx"),
                        str(tmp_path / "b.mtx")], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = reference(A, B)
    p = np.fromfile(tmp_path / "Ptr.output", dtype=np.int32).astype(np.int64)
    i = np.fromfile(tmp_path / "Idx.output", dtype=np.int32)
    v = np.fromfile(tmp_path / "Val.output", dtype=np.float64)
    compare(dbd_amd.CSR(p, i, v, A.shape), ref, np.float64, "cli A*B")


def test_two_phase_device_api_with_torch(gpu_ok):
    import torch
    A = dbd_amd.gen_rmat(12, 8, seed=12)
    B = dbd_amd.gen_rmat(12, 8, seed=13)
    ref = reference(A, B)
    dA, dB = dbd_amd.DeviceCSR.from_host(A), dbd_amd.DeviceCSR.from_host(B)
    s = torch.cuda.Stream()
    with dbd_amd.Context(block_rows=512, stream=s) as ctx, torch.cuda.stream(s):
        ctx.set_operands(dA, dB)
        info = ctx.plan()
        Cptr = torch.empty(A.shape[0] + 1, dtype=torch.int64, device="cuda")
        nnz = ctx.symbolic(Cptr)
        Cidx = torch.empty(nnz, dtype=torch.int32, device="cuda")
        Cval = torch.empty(nnz, dtype=torch.float64, device="cuda")
        ctx.numeric(Cptr, Cidx, Cval)
        s.synchronize()
        C = dbd_amd.CSR(Cptr.cpu().numpy(), Cidx.cpu().numpy(), Cval.cpu().numpy(), A.shape)
    compare(C, ref, np.float64, "two-phase")
    assert info["products"] == int(np.diff(B.ptr)[A.idx].sum())


def _light_then_heavy(dtype=np.float64):
    """A of 10,240 light rows (2 entries: 16 products each) then 10,240 heavy rows (16 entries: 128
    products) times B (8 entries per row). With 512-row blocks the block averages are 16 and 128;
    the reference criteria (mean + floor(1.28 sigma) = 143) make the light blocks CPU-preferred
    (analyze_Block 1: average <= 64) and the heavy ones 'any' (2)."""
    k, n = 30000, 40000
    A1 = dbd_amd.gen_er(10240, k, 2, seed=41, dtype=dtype)
    A2 = dbd_amd.gen_er(10240, k, 16, seed=42, dtype=dtype)
    ptr = np.concatenate([A1.ptr, A1.ptr[-1] + A2.ptr[1:]])
    A = dbd_amd.CSR(ptr, np.concatenate([A1.idx, A2.idx]), np.concatenate([A1.val, A2.val]), (20480, k))
    B = dbd_amd.gen_er(k, n, 8, seed=43, dtype=dtype)
    return A, B


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_hybrid_d2h_cpu_blocks_match_oracle(gpu_ok, dtype):
    """Hybrid mode (Z/distributeWorkload.cpp:69-144): the CPU-preferred blocks are multiplied by host
    threads into the host result, the GPU does the rest; the whole C equals the oracle's and the
    reported CPU share is the light half."""
    A, B = _light_then_heavy(dtype)
    ref = reference(A, B)
    with dbd_amd.Context(block_rows=512, dtype=dtype, hybrid=True, cpu_threads=4) as ctx:
        ctx.set_operands(A, B)
        info = ctx.plan()
        assert info["blocks_cpu_pref"] == 20 and info["blocks_any"] == 20, info
        t = ctx.run("d2h")
        C = ctx.result()
    compare(C, ref, dtype, "hybrid d2h")
    assert t["cpu_blocks"] == 20 and t["cpu_products"] == 10240 * 16 and t["cpu_s"] > 0, t
    # off by default: the GPU does every block
    with dbd_amd.Context(block_rows=512, dtype=dtype) as ctx:
        ctx.set_operands(A, B)
        t = ctx.run("d2h")
        assert t["cpu_blocks"] == 0
        compare(ctx.result(), ref, dtype, "d2h without hybrid")


def test_hybrid_io_files_and_cli_report(gpu_ok, tmp_path):
    A, B = _light_then_heavy(np.float64)
    ref = reference(A, B)
    with dbd_amd.Context(block_rows=512, hybrid=True, out_dir=str(tmp_path), chunks_hint=3) as ctx:
        ctx.set_operands(A, B)
        t = ctx.run("io")
    p = np.fromfile(tmp_path / "Ptr.output", dtype=np.int32).astype(np.int64)
    i = np.fromfile(tmp_path / "Idx.output", dtype=np.int32)
    v = np.fromfile(tmp_path / "Val.output", dtype=np.float64)
    compare(dbd_amd.CSR(p, i, v, (A.shape[0], B.shape[1])), ref, np.float64, "hybrid io")
    assert t["cpu_blocks"] == 20 and t["io_bytes"] == 4 * (A.shape[0] + 1) + 12 * i.size
    # the CLI's workload distribution lines report the CPU's share
    for name, M in (("a.mtx", A), ("b.mtx", B)):
        with open(tmp_path / name, "w") as f:
            f.write("%%MatrixMarket matrix coordinate real general\n")
            f.write(f"{M.shape[0] - 1} {M.shape[1] - 1} {M.nnz}\n")
            rows = np.repeat(np.arange(M.shape[0]), np.diff(M.ptr))
            f.write("".join(f"{r_} {c_} {float(v_):.17g}\n" for r_, c_, v_ in zip(rows, M.idx, M.val)))
    binp = os.path.join(ROOT, "dbd_amd", "bin", "SSpMM")
    r = subprocess.run([binp, "--fp64", "--no-cache", "--hybrid", "--block-rows", "512", str(tmp_path / "a.mtx"),
                        str(tmp_path / "b.mtx")], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"CPUs\tBlocks:\s+(\d+) \(([0-9.]+) %\)\n\tElements: (\d+)", r.stdout)
    assert m and int(m.group(1)) == 20 and int(m.group(3)) == 10240 * 16, r.stdout
    assert float(re.search(r"Compute CPU\t\t([0-9.]+) s", r.stdout).group(1)) > 0
    i2 = np.fromfile(tmp_path / "Idx.output", dtype=np.int32)
    assert np.array_equal(i2, i)


def test_hybrid_needs_host_operands(gpu_ok):
    A, B = _light_then_heavy(np.float64)
    dA, dB = dbd_amd.DeviceCSR.from_host(A), dbd_amd.DeviceCSR.from_host(B)
    with dbd_amd.Context(block_rows=512, hybrid=True) as ctx:
        ctx.set_operands(dA, dB)
        with pytest.raises(dbd_amd.DBDError) as e:
            ctx.run("d2h")
        assert e.value.status == 1 and "host operand" in str(e.value)
        t = ctx.run("compute")  # ignored in compute mode
        assert t["cpu_blocks"] == 0
