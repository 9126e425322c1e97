"""GPU parity: the HIP path (through the C ABI) against the oracle on the same inputs.
Structure bit-exact, values within rtol * sum|a b| (1e-12 FP64, 1e-5 FP32)."""
import json
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import dbd_amd
import oracle
from helpers import check_product, compare, gpu_spgemm, reference

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def csr_sp(m, dtype=np.float64):
    return dbd_amd.CSR.from_scipy(sp.csr_matrix(m), dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_kat_bhsparse(gpu_ok, dtype):
    kat = json.load(open(os.path.join(GOLD, "bhsparse_kat.json")))
    A = dbd_amd.CSR(np.array(kat["A"]["ptr"]), np.array(kat["A"]["idx"]), np.array(kat["A"]["val"], dtype),
                    kat["A"]["shape"])
    B = dbd_amd.CSR(np.array(kat["B"]["ptr"]), np.array(kat["B"]["idx"]), np.array(kat["B"]["val"], dtype),
                    kat["B"]["shape"])
    C, t, info = gpu_spgemm(A, B, block_rows=256)
    assert C.ptr.tolist() == kat["C"]["ptr"]
    assert C.idx.tolist() == kat["C"]["idx"]
    assert C.val.tolist() == kat["C"]["val"]
    assert info["products"] == kat["products"] and info["nnz"] == 6


def test_cage4_and_laplacian(gpu_ok):
    for name in ("cage4.mtx", "laplacian_5pt_10x10.mtx"):
        A = csr_sp(scipy.io.mmread(os.path.join(GOLD, name)))
        check_product(A, what=name, block_rows=256)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape,density,seed", [((300, 200, 250), 0.02, 1), ((1000, 1000, 1000), 0.01, 2),
                                                ((64, 5000, 3000), 0.01, 3), ((2000, 300, 4000), 0.05, 4)])
def test_random(gpu_ok, dtype, shape, density, seed):
    m, k, n = shape
    rng = np.random.default_rng(seed)
    A = csr_sp(sp.random(m, k, density=density, random_state=seed, format="csr",
                         data_rvs=lambda s: rng.standard_normal(s)), dtype)
    B = csr_sp(sp.random(k, n, density=density, random_state=seed + 7, format="csr",
                         data_rvs=lambda s: rng.standard_normal(s)), dtype)
    check_product(A, B, what=f"random{shape}", block_rows=256)


def test_empty_and_degenerate(gpu_ok):
    # all-empty A
    A = dbd_amd.CSR(np.zeros(11, np.int64), np.zeros(0, np.int32), np.zeros(0), (10, 10))
    C, t, info = gpu_spgemm(A, block_rows=256)
    assert C.nnz == 0 and C.ptr.tolist() == [0] * 11
    # A non-empty but B rows empty -> all P_i = 0
    A = csr_sp(sp.eye(300, 40, format="csr"))
    B = dbd_amd.CSR(np.zeros(41, np.int64), np.zeros(0, np.int32), np.zeros(0), (40, 70))
    C, t, info = gpu_spgemm(A, B, block_rows=256)
    assert C.nnz == 0 and info["products"] == 0
    # one row, one column
    A = csr_sp(sp.csr_matrix(np.array([[2.0]])))
    C, _, _ = gpu_spgemm(A, block_rows=256)
    assert C.val.tolist() == [4.0]
    # ragged: every other row empty, rectangular
    M = sp.random(777, 333, density=0.03, random_state=5, format="lil")
    M[::2, :] = 0
    A = csr_sp(M.tocsr())
    B = csr_sp(sp.random(333, 999, density=0.02, random_state=6, format="csr"))
    check_product(A, B, what="ragged", block_rows=256)


def _skewed_matrix(n_rows, ncols, long_rows, seed, zipf=1.6):
    """Mostly short rows plus a few very long ones, so every category (tiny hash tables up to
    the LDS window path) gets rows."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(n_rows):
        if i in long_rows:
            d = long_rows[i]
        else:
            d = int(min(ncols, rng.zipf(zipf)))
        c = rng.choice(ncols, size=min(d, ncols), replace=False)
        rows.append(np.full(c.size, i))
        cols.append(c)
    r, c = np.concatenate(rows), np.concatenate(cols)
    v = rng.standard_normal(r.size)
    return sp.csr_matrix((v, (r, c)), shape=(n_rows, ncols))


@pytest.mark.parametrize("table", [0, 1])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_all_categories(gpu_ok, dtype, table):
    """Row lengths from 1 to tens of thousands: every hash-table size and both window paths
    (symbolic key min(P_i, ncols) above 8192, numeric nnz_i above 4096) get rows. Table 0 (the
    default) routes this 60,000-column problem's hub rows to the numeric hub range kernels (4
    column ranges); table 1 (hub kernels from 2^21 columns only) keeps them in the long category."""
    assert dbd_amd.Context.table_knobs(0)["hub_min_cols"] == 0
    assert dbd_amd.Context.table_knobs(1)["hub_min_cols"] == 1 << 21
    n, ncols = 6000, 60000
    long_rows = {17: 3000, 900: 5000, 4500: 1200, 5999: 40}
    A = _skewed_matrix(n, n, long_rows, 11)
    B = _skewed_matrix(n, ncols, {3: 2500, 4: 20000, 100: 800}, 12, zipf=1.4)
    Ac, Bc = csr_sp(A, dtype), csr_sp(B, dtype)
    C, t, info = check_product(Ac, Bc, what=f"all-categories, table {table}", block_rows=512, table=table)
    # the symbolic pass's last category is the long rows; the numeric pass's the hub rows (long
    # rows of more than 1024 A entries: rows 17, 900 and 4500), the one before it the other long rows
    assert info["sym_rows"][-1] >= 4, info["sym_rows"]
    if table == 0:
        assert info["num_rows"][-1] >= 3, info["num_rows"]
    else:
        assert info["num_rows"][-1] == 0 and info["num_rows"][-2] >= 4, info["num_rows"]
    assert sum(1 for x in info["num_rows"][1:] if x > 0) >= 6, info["num_rows"]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_long_rows_wide_column_range(gpu_ok, dtype):
    """Long rows (nnz_i > 4096) of 30-36 A entries over a 5,000,000-column range: the numeric
    window sweep (dense column windows, cursors in registers) and the symbolic bitmap windows."""
    rng = np.random.default_rng(21)
    ncols = 5_000_000
    rows_b, cols_b = [], []
    for i in range(64):
        c = rng.choice(ncols, size=300, replace=False)
        rows_b.append(np.full(c.size, i))
        cols_b.append(c)
    B = sp.csr_matrix((rng.standard_normal(64 * 300), (np.concatenate(rows_b), np.concatenate(cols_b))),
                      shape=(64, ncols))
    rows_a, cols_a = [], []
    for i in range(50):
        c = rng.choice(64, size=30 + (i % 7), replace=False)
        rows_a.append(np.full(c.size, i))
        cols_a.append(c)
    A = sp.csr_matrix((rng.standard_normal(sum(x.size for x in cols_a)),
                       (np.concatenate(rows_a), np.concatenate(cols_a))), shape=(50, 64))
    C, t, info = check_product(csr_sp(A, dtype), csr_sp(B, dtype), what="wide long rows", block_rows=512)
    # long rows (numeric: the category before the hub rows), none with more than 1024 A entries
    assert info["num_rows"][-2] == 50 and info["sym_rows"][-1] == 50, (info["num_rows"], info["sym_rows"])
    assert info["num_rows"][-1] == 0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_long_rows_dense_column_band(gpu_ok, dtype):
    """Long rows whose output fills a dense band of 30,000 columns: the wide numeric windows
    hold more columns than their value slots, so every window is cut at the slot count and the
    next one narrowed (then widened again past the band)."""
    rng = np.random.default_rng(5)
    ncols = 100_000
    rows_b, cols_b = [], []
    for i in range(40):
        c = np.concatenate([rng.choice(np.arange(10_000, 40_000), 3000, replace=False),
                            rng.choice(np.concatenate([np.arange(0, 10_000), np.arange(40_000, ncols)]), 50,
                                       replace=False)])
        rows_b.append(np.full(c.size, i))
        cols_b.append(c)
    B = sp.csr_matrix((rng.standard_normal(sum(x.size for x in cols_b)),
                       (np.concatenate(rows_b), np.concatenate(cols_b))), shape=(40, ncols))
    rows_a, cols_a = [], []
    for i in range(30):
        c = rng.choice(40, size=8 + i, replace=False)
        rows_a.append(np.full(c.size, i))
        cols_a.append(c)
    A = sp.csr_matrix((rng.standard_normal(sum(x.size for x in cols_a)),
                       (np.concatenate(rows_a), np.concatenate(cols_a))), shape=(30, 40))
    C, t, info = check_product(csr_sp(A, dtype), csr_sp(B, dtype), what="dense band", block_rows=512)
    assert info["num_rows"][-2] == 30, info["num_rows"]
    assert int(np.diff(C.ptr).min()) > 8192


def test_block_rows_do_not_change_result(gpu_ok):
    A = csr_sp(_skewed_matrix(3000, 3000, {5: 900}, 21))
    ref = reference(A)
    for br in (256, 512, 4096):
        C, _, info = gpu_spgemm(A, block_rows=br)
        compare(C, ref, np.float64, f"block_rows={br}")
        assert info["num_blocks"] == (3000 - 1) // br + 1


def test_plan_matches_oracle(gpu_ok):
    """Row products, blocks and the reference criteria from the device plan vs the oracle."""
    A = csr_sp(_skewed_matrix(5000, 5000, {7: 2000}, 31))
    with dbd_amd.Context(block_rows=512) as ctx:
        ctx.set_operands(A)
        info = ctx.plan()
        P = ctx.row_products(5000)
        bptr, bprod = ctx.blocks(info["num_blocks"])
    Po = oracle.row_products(A.ptr, A.idx, A.ptr)
    assert P.tolist() == Po.tolist()
    bi = oracle.blocks(Po, 512)
    assert bprod.tolist() == bi.tolist()
    crit = oracle.criteria(bi, 512)
    assert info["criteria"] == crit and info["products"] == int(Po.sum())
    cats = [oracle.analyze_block(crit, int(b // 512)) for b in bi]
    assert (info["blocks_gpu_only"], info["blocks_cpu_pref"], info["blocks_any"]) == (
        cats.count(0), cats.count(1), cats.count(2))


def test_web_like_full_size(gpu_ok):
    """Web-like graph of 281,903 rows (Zipf out-degree, 70% of targets within +-2000 of the
    row), C = A*A in FP64 with 4096-row blocks: the CPU-runnable configuration, full size."""
    for s in (2.1, 1.66):
        A = dbd_amd.gen_web(281903, zipf_s=s, seed=1)
        check_product(A, what=f"web s={s}", block_rows=4096)


def test_rmat_fp32_a_times_b(gpu_ok):
    A = dbd_amd.gen_rmat(14, 16, seed=1, add_identity=True, row_normalize=True, dtype=np.float32)
    B = dbd_amd.gen_rmat(14, 16, seed=2, dtype=np.float32)
    check_product(A, B, what="rmat14 fp32", block_rows=512)


def test_rmat_fp64_a_times_a(gpu_ok):
    A = dbd_amd.gen_rmat(15, 16, seed=3)
    check_product(A, what="rmat15 fp64", block_rows=512)


def test_er_16_per_row_sampled_full_size(gpu_ok):
    """Erdos-Renyi 2,000,000^2 with 16 nnz/row, C = A*B FP64, 512-row blocks, 32 GB device
    budget (the benchmark configuration): full-size properties plus sampled rows vs oracle."""
    n = 2_000_000
    A = dbd_amd.gen_er(n, n, 16, seed=1)
    B = dbd_amd.gen_er(n, n, 16, seed=2)
    C, t, info = gpu_spgemm(A, B, block_rows=512, device_gib=32)
    assert info["products"] == 512_000_000
    assert C.nnz == info["nnz"] and C.ptr[-1] == C.nnz
    # sorted, in range
    d = np.diff(C.idx.astype(np.int64))
    row_start = np.zeros(C.nnz, bool)
    row_start[C.ptr[:-1][np.diff(C.ptr) > 0]] = True
    assert np.all((d > 0) | row_start[1:])
    assert C.idx.min() >= 0 and C.idx.max() < n
    # sampled rows vs oracle
    rng = np.random.default_rng(0)
    sel = np.sort(rng.choice(n, 4000, replace=False)).astype(np.int32)
    Sptr, Sidx, Sval = oracle.spgemm_rows(sel, A, B)
    for j, i in enumerate(sel):
        a, b = C.ptr[i], C.ptr[i + 1]
        assert np.array_equal(C.idx[a:b], Sidx[Sptr[j]:Sptr[j + 1]]), f"row {i}"
        np.testing.assert_allclose(C.val[a:b], Sval[Sptr[j]:Sptr[j + 1]], rtol=1e-12, atol=1e-12)
    # linearity: C x = A (B x) for random x, over the whole product
    x = rng.standard_normal(n)
    Csp = sp.csr_matrix((C.val, C.idx, C.ptr), shape=(n, n))
    Asp = sp.csr_matrix((A.val, A.idx, A.ptr), shape=(n, n))
    Bsp = sp.csr_matrix((B.val, B.idx, B.ptr), shape=(n, n))
    np.testing.assert_allclose(Csp @ x, Asp @ (Bsp @ x), rtol=1e-9, atol=1e-9)


def _mostly_distinct(m, k, n, seed, dtype):
    """A (m x k) * B (k x n) where most rows' products fall on distinct columns (B rows of 8
    random columns out of n = 10^6), plus rows that exercise the hash path inside the same
    kernel: rows with two B rows from a block of identical patterns (duplicate products), rows
    selecting B rows of 20 entries (longer than a gather step), rows of 17-24 entries."""
    rng = np.random.default_rng(seed)
    bl = np.full(k, 8)
    bl[100:200] = 20
    bptr = np.zeros(k + 1, np.int64)
    bptr[1:] = np.cumsum(bl)
    bidx = np.empty(bptr[-1], np.int32)
    for r in range(k):
        cols = np.arange(8) if r < 100 else rng.choice(n, bl[r], replace=False)
        bidx[bptr[r]:bptr[r + 1]] = np.sort(cols)
    B = dbd_amd.CSR(bptr, bidx, rng.standard_normal(bptr[-1]).astype(dtype), (k, n))
    rows = []
    for i in range(m):
        kind = rng.random()
        if kind < 0.03:
            e = np.concatenate([rng.choice(100, 2, replace=False), rng.choice(np.arange(200, k), 10, replace=False)])
        elif kind < 0.06:
            e = np.concatenate([rng.choice(np.arange(100, 200), 3, replace=False),
                                rng.choice(np.arange(200, k), 8, replace=False)])
        elif kind < 0.08:
            e = rng.choice(np.arange(200, k), int(rng.integers(17, 25)), replace=False)
        else:
            e = rng.choice(np.arange(200, k), int(rng.integers(1, 17)), replace=False)
        rows.append(np.sort(e))
    aptr = np.zeros(m + 1, np.int64)
    aptr[1:] = np.cumsum([len(e) for e in rows])
    aidx = np.concatenate(rows).astype(np.int32)
    A = dbd_amd.CSR(aptr, aidx, rng.standard_normal(aptr[-1]).astype(dtype), (m, k))
    return A, B


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_distinct_products_register_path(gpu_ok, dtype):
    """Problems with few duplicate products run the one-wave numeric shapes with the register
    path for rows of all-distinct products; rows with duplicates, long B rows or many entries
    take the hash path of the same kernel. Every candidate table (G = 16 and G = 64 shapes)."""
    A, B = _mostly_distinct(20000, 5000, 1_000_000, 77, dtype)
    ref = reference(A, B)
    for tid in range(dbd_amd.Context.table_count()):
        with dbd_amd.Context(block_rows=512, dtype=dtype) as ctx:
            ctx.set_table(tid)
            ctx.set_operands(A, B)
            ctx.run("compute")
            info = ctx.plan_info()
            assert info["nnz"] >= 0.9 * info["products"] > 0
            compare(ctx.result(), ref, dtype, f"table {tid}")


def test_autotuned_tables_give_identical_results(gpu_ok):
    """Property: every candidate category table (the autotuner's search space) gives the same
    structure and oracle-exact values; the tuned choice is one of them."""
    A = csr_sp(_skewed_matrix(4000, 4000, {1: 1500, 2: 3000}, 41))
    ref = reference(A)
    for tid in range(dbd_amd.Context.table_count()):
        with dbd_amd.Context(block_rows=512) as ctx:
            ctx.set_table(tid)
            ctx.set_operands(A)
            ctx.run("compute")
            compare(ctx.result(), ref, np.float64, f"table {tid}")
    with dbd_amd.Context(block_rows=512, tune=True) as ctx:
        ctx.set_operands(A)
        chosen = ctx.autotune()
        ctx.run("compute")
        C = ctx.result()
    assert 0 <= chosen < dbd_amd.Context.table_count()
    compare(C, ref, np.float64, f"tuned table {chosen}")
