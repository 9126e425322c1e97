"""Host-side logic of libdbd (no GPU): generators, matrix file formats, output files and
block planning vs the oracle."""
import os

import numpy as np
import pytest
import scipy.io

import dbd_amd
import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rows_sorted_unique(M):
    for i in range(M.shape[0]):
        r = M.idx[M.ptr[i]:M.ptr[i + 1]]
        if r.size and (np.any(np.diff(r) <= 0) or r[0] < 0 or r[-1] >= M.shape[1]):
            return False
    return True


def test_gen_er_exact_degree_and_determinism():
    A = dbd_amd.gen_er(5000, 7000, 16, seed=3)
    assert np.all(np.diff(A.ptr) == 16) and rows_sorted_unique(A)
    B = dbd_amd.gen_er(5000, 7000, 16, seed=3)
    assert np.array_equal(A.idx, B.idx) and np.array_equal(A.val, B.val)
    C = dbd_amd.gen_er(5000, 7000, 16, seed=4)
    assert not np.array_equal(A.idx, C.idx)
    assert abs(A.val.mean()) < 0.05 and abs(A.val.std() - 1.0) < 0.05
    # columns uniform: each decile holds ~10%
    h = np.histogram(A.idx, bins=10, range=(0, 7000))[0] / A.nnz
    assert np.all(np.abs(h - 0.1) < 0.01)
    # deg >= ncols gives full rows
    F = dbd_amd.gen_er(3, 5, 9, seed=1)
    assert F.idx.tolist() == list(range(5)) * 3
    f32 = dbd_amd.gen_er(10, 10, 3, seed=1, dtype=np.float32)
    assert f32.val.dtype == np.float32


def test_gen_web_properties():
    W = dbd_amd.gen_web(50000, zipf_s=2.1, seed=1)
    d = np.diff(W.ptr)
    assert d.min() >= 1 and d.max() <= 255 and rows_sorted_unique(W)
    # literal Zipf(2.1) on [1, 255] has mean ~3.1
    assert 2.8 < d.mean() < 3.4
    rows = np.repeat(np.arange(50000), d)
    local = np.abs(W.idx.astype(np.int64) - rows) <= 2000
    assert 0.66 < local.mean() < 0.76  # 70% local + the uniform ones that land in the band
    assert np.all((W.val >= 0) & (W.val < 1))
    W2 = dbd_amd.gen_web(50000, zipf_s=1.66, seed=1)
    assert np.diff(W2.ptr).mean() > 6.0


def test_gen_rmat_properties():
    R = dbd_amd.gen_rmat(12, 16, seed=1)
    n = 4096
    assert R.shape == (n, n) and R.nnz <= 16 * n and rows_sorted_unique(R)
    assert np.all((R.val > 0) & (R.val <= 1))
    rows = np.repeat(np.arange(n), np.diff(R.ptr))
    # a + b = 0.76 of generated edges start in the top half; merging duplicates removes more
    # of them there (the top half is denser), so the kept share is slightly lower
    assert 0.70 < (rows < n // 2).mean() < 0.80
    N = dbd_amd.gen_rmat(10, 8, seed=2, add_identity=True, row_normalize=True)
    s = np.add.reduceat(N.val, N.ptr[:-1]) if N.nnz else np.zeros(0)
    assert np.allclose(s, 1.0)
    for i in range(1024):
        assert i in N.idx[N.ptr[i]:N.ptr[i + 1]]
    assert np.array_equal(dbd_amd.gen_rmat(10, 8, seed=5).idx, dbd_amd.gen_rmat(10, 8, seed=5).idx)


def test_gen_skewed_mean_and_skew():
    S = dbd_amd.gen_skewed(20000, 32.0, zipf_s=1.8, seed=1)
    d = np.diff(S.ptr)
    assert 28 < d.mean() < 36 and d.max() > 10 * d.mean() and rows_sorted_unique(S)


def test_read_mtx_reference_conventions():
    """cage4.mtx through make_CSR's conventions: dims = header + 1 (1-based indices kept),
    row/column 0 empty, values as in the file."""
    M = dbd_amd.read_matrix(os.path.join(GOLD, "cage4.mtx"), dtype=np.float64)
    S = scipy.io.mmread(os.path.join(GOLD, "cage4.mtx")).tocsr()
    assert M.shape == (10, 10) and M.nnz == 49
    assert M.ptr[1] == 0  # row 0 empty
    D = np.zeros((10, 10))
    D[1:, 1:] = S.toarray()
    np.testing.assert_array_equal(M.to_dense(), D)
    assert rows_sorted_unique(M)


def test_read_mtx_symmetric_pattern():
    M = dbd_amd.read_matrix(os.path.join(GOLD, "coordinate_pattern_symmetric.mtx"), dtype=np.float32)
    assert M.shape == (6, 6)
    D = M.to_dense()
    assert np.array_equal(D, D.T) and set(np.unique(D)) <= {0.0, 1.0}
    assert D[4, 2] == 1 and D[2, 4] == 1 and M.nnz == 9  # 7 entries, 2 off-diagonal mirrored


def test_csr_cache_round_trip(tmp_path):
    src = tmp_path / "m.mtx"
    src.write_text(open(os.path.join(GOLD, "coordinate_real_general.mtx")).read())
    M = dbd_amd.read_matrix(str(src), dtype=np.float32, use_cache=True)
    assert (tmp_path / "m.mtx.csr").exists()
    raw = np.fromfile(tmp_path / "m.mtx.csr", dtype=np.int32)
    assert raw[:3].tolist() == [6, 6, 8]
    M2 = dbd_amd.read_matrix(str(src), dtype=np.float32, use_cache=True)  # from the cache now
    assert np.array_equal(M.ptr, M2.ptr) and np.array_equal(M.idx, M2.idx) and np.array_equal(M.val, M2.val)
    assert M.val.tolist()[:2] == [1.0, 6.0]


def test_fp64_cache_keeps_fp64_values(tmp_path):
    """An FP64 read with the cache on must give the same values the second time (from the
    cache) as the first (from the text): the FP64 cache is <file>.csr64 with double values, and a
    float .csr cache written by an FP32 run is not used for FP64."""
    src = tmp_path / "m.mtx"
    src.write_text("%%MatrixMarket matrix coordinate real general\n3 3 4\n"
                   "1 1 0.1\n1 3 1.000000001\n2 2 3.141592653589793\n3 1 -2.718281828459045\n")
    M32 = dbd_amd.read_matrix(str(src), dtype=np.float32, use_cache=True)  # writes the float .csr
    assert (tmp_path / "m.mtx.csr").exists()
    M1 = dbd_amd.read_matrix(str(src), dtype=np.float64, use_cache=True)
    assert (tmp_path / "m.mtx.csr64").exists()
    M2 = dbd_amd.read_matrix(str(src), dtype=np.float64, use_cache=True)  # from .csr64 now
    assert M1.val.tolist() == [0.1, 1.000000001, 3.141592653589793, -2.718281828459045]
    assert np.array_equal(M1.ptr, M2.ptr) and np.array_equal(M1.idx, M2.idx)
    assert M1.val.tolist() == M2.val.tolist()
    assert M32.val.tolist() == np.float32([0.1, 1.000000001, 3.141592653589793, -2.718281828459045]).tolist()
    raw = np.fromfile(tmp_path / "m.mtx.csr64", dtype=np.int32)
    assert raw[:3].tolist() == [4, 4, 4]


def test_read_dat_without_banner(tmp_path):
    """make_CSR (Z/readInput.cpp:52-56) reads the first line of a .dat file as the size line:
    there is no MatrixMarket banner to skip. The same entries as a .mtx give the same matrix."""
    body = "3 4 5\n1 1 2.0\n1 4 1.5\n2 2 -1.0\n3 1 4.0\n3 3 0.5\n"
    dat = tmp_path / "m.dat"
    dat.write_text(body)
    mtx = tmp_path / "m.mtx"
    mtx.write_text("%%MatrixMarket matrix coordinate real general\n" + body)
    D = dbd_amd.read_matrix(str(dat), dtype=np.float64)
    M = dbd_amd.read_matrix(str(mtx), dtype=np.float64)
    assert D.shape == (4, 5) == M.shape
    assert np.array_equal(D.ptr, M.ptr) and np.array_equal(D.idx, M.idx) and np.array_equal(D.val, M.val)
    dense = np.zeros((4, 5))
    dense[1, 1], dense[1, 4], dense[2, 2], dense[3, 1], dense[3, 3] = 2.0, 1.5, -1.0, 4.0, 0.5
    np.testing.assert_array_equal(D.to_dense(), dense)
    # a .mtx without its banner loses the size line to the skip: the reader refuses it
    bad = tmp_path / "nobanner.mtx"
    bad.write_text(body)
    with pytest.raises(dbd_amd.DBDError):
        dbd_amd.read_matrix(str(bad), dtype=np.float64)


def test_read_errors(tmp_path):
    with pytest.raises(dbd_amd.DBDError) as e:
        dbd_amd.read_matrix(str(tmp_path / "missing.mtx"))
    assert e.value.status == 5
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate real general\n3\n")
    with pytest.raises(dbd_amd.DBDError) as e:
        dbd_amd.read_matrix(str(bad))
    assert e.value.status == 1
    out = tmp_path / "oob.mtx"
    out.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n5 1 1.0\n")
    with pytest.raises(dbd_amd.DBDError):
        dbd_amd.read_matrix(str(out))


def test_write_outputs_formats(tmp_path):
    A = dbd_amd.gen_er(100, 50, 4, seed=1)
    dbd_amd.write_outputs(str(tmp_path), A)
    p = np.fromfile(tmp_path / "Ptr.output", dtype=np.int32)
    assert p.tolist() == A.ptr.tolist()
    assert np.fromfile(tmp_path / "Idx.output", dtype=np.int32).tolist() == A.idx.tolist()
    assert np.array_equal(np.fromfile(tmp_path / "Val.output", dtype=np.float64), A.val)
    dbd_amd.write_outputs(str(tmp_path), A.astype(np.float32), ptr64=True)
    assert np.fromfile(tmp_path / "Ptr.output", dtype=np.int64).tolist() == A.ptr.tolist()
    assert np.fromfile(tmp_path / "Val.output", dtype=np.float32).size == A.nnz
    with pytest.raises(dbd_amd.DBDError):
        dbd_amd.write_outputs(str(tmp_path / "no" / "such"), A)


@pytest.mark.parametrize("nr,br,seed", [(5000, 512, 1), (4096 * 3, 4096, 2), (1, 256, 3), (70000, 4096, 4)])
def test_plan_blocks_vs_oracle(nr, br, seed):
    rng = np.random.default_rng(seed)
    P = rng.integers(0, 400, nr).astype(np.int64) * rng.integers(0, 3, nr)
    bptr, bprod, info = dbd_amd.plan_blocks(P, br)
    bi = oracle.blocks(P, br)
    assert bprod.tolist() == bi.tolist()
    assert bptr[0] == 0 and bptr[-1] == nr and len(bptr) == len(bi) + 1
    c = oracle.criteria(bi, br)
    assert info["criteria"] == c and info["products"] == int(P.sum())
    cats = [oracle.analyze_block(c, int(b // br)) for b in bi]
    assert (info["blocks_gpu_only"], info["blocks_cpu_pref"], info["blocks_any"]) == (
        cats.count(0), cats.count(1), cats.count(2))
    for avg in (0, 64, 65, c, c + 1):
        assert dbd_amd.analyze_block(c, avg) == oracle.analyze_block(c, avg)


def test_batch_concat_split_round_trip():
    """The batched form's block-diagonal layout (host logic only): the product of the
    concatenations, split per member, equals every member's own product (oracle on CPU)."""
    from dbd_amd import batch
    rng = np.random.default_rng(1)
    As, Bs = [], []
    for i in range(5):
        n = int(np.exp(rng.uniform(np.log(200), np.log(6000))))
        m = float(rng.uniform(4, 64))
        As.append(dbd_amd.gen_skewed(n, m, 1.8, seed=1000 + 2 * i, dtype=np.float32))
        Bs.append(dbd_amd.gen_skewed(n, m, 1.8, seed=1000 + 2 * i + 1, dtype=np.float32))
    A_cat, B_cat, lay = batch.concat(As, Bs)
    assert A_cat.shape[0] == sum(A.shape[0] for A in As)
    Cp, Ci, Cv = oracle.spgemm(A_cat, B_cat)
    parts = batch.split(Cp, Ci, Cv, lay)
    for A, B, Cm in zip(As, Bs, parts):
        p, i, v = oracle.spgemm(A, B)
        assert np.array_equal(Cm.ptr, p) and np.array_equal(Cm.idx, i) and np.allclose(Cm.val, v)
