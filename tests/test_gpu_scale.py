"""Parity at the named configurations' own sizes, where 64-bit counts and out-of-core chunks
are exercised: every product count, row pointer and chunk offset that passes 2^31 is read back
and checked against the oracle.

  - R-MAT scale 22 (a,b,c,d = .57,.19,.19,.05) C = A*A, FP64, 512-row blocks, 64 GB device
    budget, on the 16 heaviest hub rows plus the trailing 2^19 rows: P > 2^31 products;
    every row compared with the oracle.
  - Erdos-Renyi 2,000,000^2 with 33 nnz per row, C = A*B, FP64, 512-row blocks, 32 GB device
    budget, written to files (the reference's SSpMM mode): nnz(C) > 2^31, so Ptr.output is int64
    and the chunk offsets pass int32; the files are compared with the oracle.
  - R-MAT scale 18 forward + backward through torch.autograd, FP32 (the differentiated
    configuration at its size, P = 2.93e9): central finite differences (h = 1e-3, FP64) on 64
    random entries of A and of B, and the FP64 oracle gradient on sampled rows of A and B.
  - The 512-member batch (FP32, rows log-uniform in [1e3, 2e5], nnz per row U[4, 64] with
    Zipf(1.8) row skew): every member equals the plain per-matrix pipeline, a sample of members
    equals the oracle.

Comparisons run over row ranges so host memory stays bounded.
"""
import os
import shutil
import tempfile

import numpy as np
import pytest

import dbd_amd
import oracle
from helpers import RTOL

pytestmark = pytest.mark.gpu
INT32_MAX = 2 ** 31 - 1


def threads():
    t = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    return max(1, min(t, os.cpu_count() or 1, 16))


def compare_rows(rows, Cptr, Cidx, Cval, A, B, dtype, what, step=1 << 17):
    """Rows `rows` of the GPU result (Cptr over those rows, Cidx/Cval indexable by its offsets)
    against the oracle, in slices of `step` rows: structure exact, values within
    rtol * sum |a_ik b_kj| (the bound of any summation order)."""
    rtol = RTOL[np.dtype(dtype)]
    for s0 in range(0, len(rows), step):
        sel = np.ascontiguousarray(rows[s0:s0 + step], dtype=np.int32)
        Sp, Si, Sv, Sa = oracle.spgemm_rows(sel, A, B, threads=threads(), with_abs=True)
        e0, e1 = int(Cptr[s0]), int(Cptr[s0 + len(sel)])
        gp = Cptr[s0:s0 + len(sel) + 1] - e0
        assert np.array_equal(gp, Sp), f"{what}: row pointer differs in rows {sel[0]}.."
        gi = np.asarray(Cidx[e0:e1])
        assert np.array_equal(gi, Si), f"{what}: column indices differ in rows {sel[0]}.."
        gv = np.asarray(Cval[e0:e1]).astype(np.float64)
        bad = np.abs(gv - Sv) > rtol * Sa + 1e-300
        assert not bad.any(), f"{what}: {int(bad.sum())} values off in rows {sel[0]}.."
        print(f"{what}: rows {s0}..{s0 + len(sel)} of {len(rows)} match ({e1 - e0} entries)", flush=True)


def row_subset(A, rows):
    """CSR of the given rows of A (in that order), same columns."""
    rows = np.asarray(rows, dtype=np.int64)
    lens = A.ptr[rows + 1] - A.ptr[rows]
    ptr = np.zeros(len(rows) + 1, np.int64)
    ptr[1:] = np.cumsum(lens)
    take = np.concatenate([np.arange(A.ptr[r], A.ptr[r + 1]) for r in rows]) if len(rows) else np.zeros(0, np.int64)
    return dbd_amd.CSR(ptr, A.idx[take], A.val[take], (len(rows), A.shape[1]))


def test_rmat22_hub_rows_and_trailing_rows_products_above_2_31(gpu_ok):
    A = dbd_amd.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=1)
    n = A.shape[0]
    P = oracle.row_products(A.ptr, A.idx, A.ptr)
    hubs = np.sort(np.argsort(P)[::-1][:16])
    trailing = np.arange(n - (1 << 19), n)
    rows = np.concatenate([hubs, trailing])
    S = row_subset(A, rows)
    Ps = int(P[rows].sum())
    assert Ps > INT32_MAX  # the product prefix passes int32
    with dbd_amd.Context(block_rows=512, dtype=np.float64, device_gib=64, host_gib=64) as ctx:
        ctx.set_operands(S, A)
        info = ctx.plan()
        assert info["products"] == Ps
        t = ctx.run("compute")
        info = ctx.plan_info()
        # 4,194,304 columns: the heaviest rows (tens of thousands of A entries) take the hub kernels
        assert info["num_rows"][-1] >= 16, info["num_rows"]
        C = ctx.result()
    print(f"R-MAT 22 slice: {len(rows)} rows, P = {Ps}, nnz(C) = {C.nnz}, kernels {t['compute_gpu_s']:.3f} s",
          flush=True)
    assert np.all(np.diff(C.ptr) >= 0) and C.ptr[-1] == C.nnz
    compare_rows(rows, C.ptr, C.idx, C.val, A, A, np.float64, "R-MAT 22 slice")


def test_hub_rows_beyond_one_launch_of_range_workgroups(gpu_ok):
    """1100 hub rows of 1030 A entries over 4096 column ranges (B: 5000 rows of 64 entries spread
    over 67,108,864 columns): 4.5 million (row, range) workgroups of 1024 threads, more than one
    launch can hold (fewer than 2^32 work-items), so the range kernel goes out in row groups.
    Sampled rows against the oracle; every row length against its product count."""
    rng = np.random.default_rng(77)
    ncols, nb, lb, nh, la = 4096 * 16384, 5000, 64, 1100, 1030
    bidx = np.empty((nb, lb), np.int32)
    for i in range(nb):
        bidx[i] = np.unique(rng.integers(0, ncols, 2 * lb))[:lb]
    B = dbd_amd.CSR(np.arange(0, nb * lb + 1, lb, dtype=np.int64), bidx.ravel(), rng.standard_normal(nb * lb),
                    (nb, ncols))
    aidx = np.stack([np.sort(rng.choice(nb, la, replace=False)) for _ in range(nh)]).astype(np.int32)
    A = dbd_amd.CSR(np.arange(0, nh * la + 1, la, dtype=np.int64), aidx.ravel(), rng.standard_normal(nh * la),
                    (nh, nb))
    with dbd_amd.Context(block_rows=512, dtype=np.float64) as ctx:
        ctx.set_operands(A, B)
        ctx.run("compute")
        info = ctx.plan_info()
        assert info["num_rows"][-1] == nh, info["num_rows"]  # every row in the hub category
        C = ctx.result()
    assert C.ptr[0] == 0 and C.ptr[-1] == C.nnz and np.all(np.diff(C.ptr) > 4096)
    assert np.all(np.diff(C.ptr) <= la * lb)
    sel = np.sort(rng.choice(nh, 24, replace=False)).astype(np.int32)
    lens = C.ptr[sel + 1] - C.ptr[sel]
    sptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    take = np.concatenate([np.arange(C.ptr[r], C.ptr[r + 1]) for r in sel])
    compare_rows(sel, sptr, C.idx[take], C.val[take], A, B, np.float64, "hub rows in row groups")


def test_er2m_33_per_row_files_nnz_above_2_31(gpu_ok):
    n, deg = 2_000_000, 33
    A = dbd_amd.gen_er(n, n, deg, seed=31)
    B = dbd_amd.gen_er(n, n, deg, seed=32)
    d = tempfile.mkdtemp(prefix="dbd_scale_io_", dir="/tmp")
    try:
        with dbd_amd.Context(block_rows=512, dtype=np.float64, device_gib=32, host_gib=48, out_dir=d,
                             writer_threads=4, chunks_hint=8) as ctx:
            ctx.set_operands(A, B)
            t = ctx.run("io")
            nnz = ctx.plan_info()["nnz"]
        print(f"ER 2M x 33: nnz(C) = {nnz}, chunks {t['chunks']}, total {t['total_s']:.2f} s", flush=True)
        assert nnz > INT32_MAX and t["chunks"] > 1
        # nnz(C) >= 2^31: Ptr.output is int64 (the reference's int32 cannot hold it)
        assert os.path.getsize(os.path.join(d, "Ptr.output")) == 8 * (n + 1)
        assert os.path.getsize(os.path.join(d, "Idx.output")) == 4 * nnz
        assert os.path.getsize(os.path.join(d, "Val.output")) == 8 * nnz
        Cptr = np.fromfile(os.path.join(d, "Ptr.output"), dtype=np.int64)
        assert Cptr[0] == 0 and Cptr[-1] == nnz and np.all(np.diff(Cptr) >= 1)
        Cidx = np.memmap(os.path.join(d, "Idx.output"), dtype=np.int32, mode="r")
        Cval = np.memmap(os.path.join(d, "Val.output"), dtype=np.float64, mode="r")
        compare_rows(np.arange(n), Cptr, Cidx, Cval, A, B, np.float64, "ER 33/row files", step=250_000)
        del Cidx, Cval
    finally:
        shutil.rmtree(d, ignore_errors=True)


def test_rmat18_forward_backward_finite_differences(gpu_ok):
    import torch
    from dbd_amd.autograd import spgemm_values
    A = dbd_amd.gen_rmat(18, 16, seed=1, add_identity=True, row_normalize=True, dtype=np.float32)
    B = dbd_amd.gen_rmat(18, 16, seed=2, dtype=np.float32)
    Ad, Bd = dbd_amd.DeviceCSR.from_host(A), dbd_amd.DeviceCSR.from_host(B)
    av = Ad.val.clone().requires_grad_(True)
    bv = Bd.val.clone().requires_grad_(True)
    Cv, Cp, Ci = spgemm_values(Ad.ptr, Ad.idx, av, A.shape, Bd.ptr, Bd.idx, bv, B.shape, block_rows=512)
    Prow = oracle.row_products(A.ptr, A.idx, B.ptr)
    P = int(Prow.sum())
    assert P > INT32_MAX
    g = torch.Generator(device="cuda").manual_seed(5)
    W = torch.randn(Cv.shape, dtype=torch.float32, device="cuda", generator=g)
    (Cv * W).sum().backward()
    gA = av.grad.cpu().numpy().astype(np.float64)
    gB = bv.grad.cpu().numpy().astype(np.float64)
    Cptr = Cp.cpu().numpy()
    Cidx = Ci.cpu().numpy()
    Wh = W.cpu().numpy()
    del Cv, Ci, W
    print(f"R-MAT 18 fwd+bwd: P = {P}, nnz(C) = {Cidx.size}", flush=True)

    def row_of(ptr, j):
        return int(np.searchsorted(ptr, j, side="right") - 1)

    def c_row(i):
        return Cidx[Cptr[i]:Cptr[i + 1]], Wh[Cptr[i]:Cptr[i + 1]].astype(np.float64)

    def rows_csr(rows):
        cp = np.zeros(len(rows) + 1, np.int64)
        cp[1:] = np.cumsum(Cptr[rows + 1] - Cptr[rows])
        take = np.concatenate([np.arange(Cptr[i], Cptr[i + 1]) for i in rows])
        return cp, Cidx[take], Wh[take].astype(np.float64)

    def fd(rows, A_pm, B_pm):
        """Central difference of L = sum W * C over the rows that move, by the FP64 oracle on
        (A_pm[0], B_pm[0]) and (A_pm[1], B_pm[1]); C's pattern must be the GPU's. Entries the
        perturbation does not touch come out bitwise equal in both runs and cancel exactly."""
        rows = np.asarray(rows, np.int32)
        outs = []
        for Am, Bm in zip(A_pm, B_pm):
            Sp, Si, Sv = oracle.spgemm_rows(rows, Am, Bm, threads=threads())
            cp, ci, wi = rows_csr(rows.astype(np.int64))
            assert np.array_equal(Sp, cp) and np.array_equal(Si, ci)
            outs.append(Sv)
        return float(np.dot(wi, outs[0] - outs[1])) / (2 * h)

    A64 = A.astype(np.float64)
    B64 = B.astype(np.float64)
    h = 1e-3
    rng = np.random.default_rng(7)
    # dA: the loss is linear in A_ik and only row i of C moves
    for j in rng.choice(A.nnz, 64, replace=False):
        i = row_of(A.ptr, j)
        Ap, Am = (dbd_amd.CSR(A64.ptr, A64.idx, A64.val.copy(), A.shape) for _ in range(2))
        Ap.val[j] += h
        Am.val[j] -= h
        d = fd([i], (Ap, Am), (B64, B64))
        ci, wi = c_row(i)
        k = A.idx[j]
        bk = slice(B.ptr[k], B.ptr[k + 1])
        scale = float(np.abs(wi[np.searchsorted(ci, B.idx[bk])]) @ np.abs(B64.val[bk]))
        assert abs(gA[j] - d) <= 1e-5 * scale + 1e-12, f"dA[{j}]: gpu {gA[j]} fd {d} scale {scale}"
    # dB: perturbing B_kj moves row i of C for every i with A_ik != 0; entries are drawn until 64
    # have affected rows of at most 2e7 products (the FD recomputes those rows twice)
    AT = A.to_scipy().tocsc()
    AT.sort_indices()
    done = 0
    for j in rng.permutation(B.nnz):
        if done == 64:
            break
        k = row_of(B.ptr, j)
        rows = AT.indices[AT.indptr[k]:AT.indptr[k + 1]]
        if rows.size == 0 or Prow[rows].sum() > 2e7:
            continue
        Bp, Bm = (dbd_amd.CSR(B64.ptr, B64.idx, B64.val.copy(), B.shape) for _ in range(2))
        Bp.val[j] += h
        Bm.val[j] -= h
        d = fd(rows, (A64, A64), (Bp, Bm))
        col = B.idx[j]
        scale = 0.0
        for i in rows:
            ci, wi = c_row(i)
            aik = A64.val[A.ptr[i] + np.searchsorted(A.idx[A.ptr[i]:A.ptr[i + 1]], k)]
            scale += abs(aik) * abs(wi[np.searchsorted(ci, col)])
        assert abs(gB[j] - d) <= 1e-5 * scale + 1e-12, f"dB[{j}]: gpu {gB[j]} fd {d} scale {scale}"
        done += 1
    assert done == 64
    # FP64 oracle gradient: dA on 2,000 sampled rows of A; dB on the two B rows that the most rows
    # of A reach within 1e8 products of those rows (the heaviest atomics the oracle can afford)
    # and 100 sampled B rows, each from every row of A that reaches it
    sel = np.sort(rng.choice(A.shape[0], 2000, replace=False))
    absB = dbd_amd.CSR(B64.ptr, B64.idx, np.abs(B64.val), B.shape)

    def grads(rows):
        Asub = row_subset(A64, rows)
        cp, ci, wi = rows_csr(rows)
        dA_ref, dB_ref, miss = oracle.backward(Asub, B64, cp, ci, wi)
        assert miss == 0
        sA, sB, _ = oracle.backward(dbd_amd.CSR(Asub.ptr, Asub.idx, np.abs(Asub.val), Asub.shape), absB, cp, ci,
                                    np.abs(wi))
        return dA_ref, dB_ref, sA, sB

    dA_ref, _, sA, _ = grads(sel)
    take = np.concatenate([np.arange(A.ptr[i], A.ptr[i + 1]) for i in sel])
    assert not (np.abs(gA[take] - dA_ref) > 1e-5 * sA + 1e-12).any(), "dA vs FP64 oracle"
    indeg = np.diff(AT.indptr)
    reach = np.zeros(B.shape[0], np.int64)  # products of the rows of A that reach each B row
    nzc = indeg > 0
    reach[nzc] = np.add.reduceat(Prow[AT.indices], AT.indptr[:-1][nzc])
    fits = np.nonzero((reach <= 1e8) & (np.diff(B.ptr) > 0))[0]
    heavy = list(fits[np.argsort(indeg[fits])[::-1][:2]])
    assert len(heavy) == 2 and indeg[heavy[1]] > 100
    ks = heavy + list(rng.choice(fits, 100, replace=False))
    for k in ks:
        rows = np.sort(AT.indices[AT.indptr[k]:AT.indptr[k + 1]]).astype(np.int64)
        if rows.size == 0:
            continue
        _, dB_ref, _, sB = grads(rows)
        bk = slice(B.ptr[k], B.ptr[k + 1])
        assert not (np.abs(gB[bk] - dB_ref[bk]) > 1e-5 * sB[bk] + 1e-12).any(), f"dB row {k} vs FP64 oracle"


def test_batch_512_members_full_size(gpu_ok):
    import torch
    from dbd_amd import batch
    rng = np.random.default_rng(2024)
    As, Bs = [], []
    for i in range(512):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(200000))))
        m = float(rng.uniform(4, 64))
        As.append(dbd_amd.gen_skewed(n, m, 1.8, seed=2024 * 10007 + 2 * i, dtype=np.float32))
        Bs.append(dbd_amd.gen_skewed(n, m, 1.8, seed=2024 * 10007 + 2 * i + 1, dtype=np.float32))
    A_cat, B_cat, lay = batch.concat(As, Bs)
    runner = batch.BatchRunner(A_cat, B_cat, block_rows=4096)
    del A_cat, B_cat
    runner.launch()
    torch.cuda.synchronize()
    nnz = runner.nnz
    P = sum(int(np.diff(B.ptr)[A.idx].sum()) for A, B in zip(As, Bs))
    print(f"batch of 512: P = {P}, nnz(C) = {nnz}", flush=True)
    assert int(runner.Cptr[-1].item()) == nnz and nnz <= P
    # every member equals the plain pipeline (on the device); a sample also equals the oracle
    ctx = dbd_amd.Context(block_rows=4096, dtype=np.float32)
    sample = set(rng.choice(512, 12, replace=False).tolist())
    for m_, (A, B) in enumerate(zip(As, Bs)):
        r0, r1 = int(lay.row_off[m_]), int(lay.row_off[m_ + 1])
        dA, dB = dbd_amd.DeviceCSR.from_host(A), dbd_amd.DeviceCSR.from_host(B)
        ctx.set_operands(dA, dB)
        ctx.plan()
        Cptr = torch.empty(A.shape[0] + 1, dtype=torch.int64, device="cuda")
        k = ctx.symbolic(Cptr)
        Cidx = torch.empty(max(k, 1), dtype=torch.int32, device="cuda")
        Cval = torch.empty(max(k, 1), dtype=torch.float32, device="cuda")
        ctx.numeric(Cptr, Cidx, Cval)
        bp = runner.Cptr[r0:r1 + 1]
        e0 = int(bp[0].item())
        assert torch.equal(bp - e0, Cptr), f"member {m_}: row pointer"
        bi = runner.Cidx[e0:e0 + k] - int(lay.col_off[m_])
        assert torch.equal(bi, Cidx[:k]), f"member {m_}: columns"
        bv = runner.Cval[e0:e0 + k]
        # same kernels and categories per row; summation order may differ only through atomics
        assert torch.allclose(bv, Cval[:k], rtol=1e-5, atol=1e-6), f"member {m_}: values"
        if m_ in sample:
            Sp, Si, Sv, Sa = oracle.spgemm(A, B, threads=threads(), with_abs=True)
            assert np.array_equal(Cptr.cpu().numpy(), Sp) and np.array_equal(Cidx[:k].cpu().numpy(), Si)
            gv = bv.cpu().numpy().astype(np.float64)
            assert not (np.abs(gv - Sv) > 1e-5 * Sa + 1e-300).any(), f"member {m_}: values vs oracle"
        del Cptr, Cidx, Cval, dA, dB
    ctx.close()
    runner.close()
