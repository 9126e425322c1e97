"""Pinning the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle function is checked against something that does not restate its code:
known-answer vectors from the reference repository's own tests, closed forms, brute force
with numpy dense products, scipy, and hand-worked numbers. A plausible mistake (dropped
product, wrong index, transposed operand, unsorted output, value-based first touch) fails at
least one of these.
"""
import json
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


class M:
    def __init__(self, ptr, idx, val, shape):
        self.ptr = np.asarray(ptr, np.int64)
        self.idx = np.asarray(idx, np.int32)
        self.val = np.asarray(val, np.float64)
        self.shape = tuple(shape)


def from_sp(m):
    m = sp.csr_matrix(m)
    m.sort_indices()
    m.sum_duplicates()
    return M(m.indptr, m.indices, m.data, m.shape)


def dense_of(ptr, idx, val, shape):
    d = np.zeros(shape)
    for i in range(shape[0]):
        for p in range(ptr[i], ptr[i + 1]):
            d[i, idx[p]] += val[p]
    return d


def pattern_product(A, B):
    """Structure of A*B by brute-force set unions (keeps entries whose value cancels)."""
    rows = []
    for i in range(A.shape[0]):
        s = set()
        for p in range(A.ptr[i], A.ptr[i + 1]):
            k = A.idx[p]
            s.update(B.idx[B.ptr[k]:B.ptr[k + 1]].tolist())
        rows.append(sorted(s))
    return rows


def check_sorted_unique(Cptr, Cidx):
    for i in range(len(Cptr) - 1):
        r = Cidx[Cptr[i]:Cptr[i + 1]]
        assert np.all(np.diff(r) > 0), f"row {i} not strictly increasing"


# ---- known-answer tests ------------------------------------------------------------------
def test_kat_bhsparse():
    """bhsparse/main.cu:182-279 small SpGEMM; expected C derived by hand (exact)."""
    kat = json.load(open(os.path.join(GOLD, "bhsparse_kat.json")))
    A = M(kat["A"]["ptr"], kat["A"]["idx"], kat["A"]["val"], kat["A"]["shape"])
    B = M(kat["B"]["ptr"], kat["B"]["idx"], kat["B"]["val"], kat["B"]["shape"])
    Cptr, Cidx, Cval = oracle.spgemm(A, B)
    assert Cptr.tolist() == kat["C"]["ptr"]
    assert Cidx.tolist() == kat["C"]["idx"]
    assert Cval.tolist() == kat["C"]["val"]
    P = oracle.row_products(A.ptr, A.idx, B.ptr)
    assert P.tolist() == kat["row_products"] and P.sum() == kat["products"]


def _mm(path):
    m = scipy.io.mmread(path).tocsr()
    return from_sp(m)


def test_cage4_squared_vs_dense():
    """bhsparse/cage4.mtx (9x9, 49 nnz): C = A*A against the numpy dense product."""
    A = _mm(os.path.join(GOLD, "cage4.mtx"))
    assert A.shape == (9, 9) and A.ptr[-1] == 49
    Cptr, Cidx, Cval = oracle.spgemm(A, A)
    D = A.val.size and dense_of(A.ptr, A.idx, A.val, A.shape)
    ref = D @ D
    got = dense_of(Cptr, Cidx, Cval, (9, 9))
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=0)
    # structure: all-positive matrix, so pattern == nonzeros of |A|@|A|
    pat = (np.abs(D) @ np.abs(D)) > 0
    assert int(pat.sum()) == Cptr[-1]
    check_sorted_unique(Cptr, Cidx)


def _cusp_dense_set():
    """The dense matrices of library/cusplibrary/testing/multiply.cu:39-97 (A..F)."""
    A = np.array([[1, 2], [3, 0], [5, 6]], float)
    B = np.array([[0, 2, 3, 4], [5, 0, 0, 8]], float)
    Cm = np.array([[0, 0], [3, 5]], float)
    D = np.array([[2], [3]], float)
    E = np.zeros((2, 2))
    F = np.array([[0, 1.5, 3], [0.5, 0, 0]], float)
    return [A, B, Cm, D, E, F]


def test_cusp_dense_pairs():
    """Every compatible pair of CUSP's test matrices: sparse (zeros dropped, as CUSP converts)
    product equals the dense product exactly."""
    mats = _cusp_dense_set()
    n = 0
    for L in mats:
        for R in mats:
            if L.shape[1] != R.shape[0]:
                continue
            Ls, Rs = from_sp(sp.csr_matrix(L)), from_sp(sp.csr_matrix(R))
            Cptr, Cidx, Cval = oracle.spgemm(Ls, Rs)
            got = dense_of(Cptr, Cidx, Cval, (L.shape[0], R.shape[1]))
            np.testing.assert_array_equal(got, L @ R)
            n += 1
    assert n >= 10


def test_poisson5pt_squared_closed_form():
    """CUSP 5-point Laplacian (library/cusplibrary/testing/data/laplacian/5pt_10x10.mtx) squared.
    Closed form of L^2 on a grid: center 16 + deg(p) (deg = grid neighbours), axis +-1: -8,
    axis +-2: 1, diagonal: 2, nothing else."""
    L = _mm(os.path.join(GOLD, "laplacian_5pt_10x10.mtx"))
    nx = ny = 10
    Cptr, Cidx, Cval = oracle.spgemm(L, L)
    got = {}
    for i in range(100):
        for p in range(Cptr[i], Cptr[i + 1]):
            got[(i, int(Cidx[p]))] = Cval[p]
    exp = {}
    inside = lambda x, y: 0 <= x < nx and 0 <= y < ny
    for y in range(ny):
        for x in range(nx):
            p = y * nx + x
            deg = sum(inside(x + dx, y + dy) for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)))
            exp[(p, p)] = 16 + deg
            for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                if inside(x + dx, y + dy):
                    exp[(p, (y + dy) * nx + x + dx)] = -8
                if inside(x + 2 * dx, y + 2 * dy):
                    exp[(p, (y + 2 * dy) * nx + x + 2 * dx)] = 1
            for dx, dy in ((1, 1), (1, -1), (-1, 1), (-1, -1)):
                if inside(x + dx, y + dy):
                    exp[(p, (y + dy) * nx + x + dx)] = 2
    assert got == exp
    check_sorted_unique(Cptr, Cidx)


def test_permutation_and_identity():
    """C = P*B for a permutation P is B's rows permuted (exact); I*B = B."""
    rng = np.random.default_rng(3)
    B = from_sp(sp.random(50, 40, density=0.1, random_state=4, format="csr"))
    perm = rng.permutation(50)
    P = M(np.arange(51), perm, np.ones(50), (50, 50))
    Cptr, Cidx, Cval = oracle.spgemm(P, B)
    for i in range(50):
        k = perm[i]
        assert Cidx[Cptr[i]:Cptr[i + 1]].tolist() == B.idx[B.ptr[k]:B.ptr[k + 1]].tolist()
        assert Cval[Cptr[i]:Cptr[i + 1]].tolist() == B.val[B.ptr[k]:B.ptr[k + 1]].tolist()
    I = M(np.arange(41), np.arange(40), np.ones(40), (40, 40))
    Cptr, Cidx, Cval = oracle.spgemm(B, I)
    assert Cptr.tolist() == B.ptr.tolist() and Cidx.tolist() == B.idx.tolist() and Cval.tolist() == B.val.tolist()


@pytest.mark.parametrize("shape,density,seed", [((30, 20, 25), 0.2, 1), ((64, 64, 64), 0.05, 2),
                                                ((7, 200, 3), 0.3, 3), ((120, 90, 300), 0.02, 4)])
def test_random_vs_bruteforce_and_scipy(shape, density, seed):
    """Structure by brute-force set unions (so cancelling sums stay structural), values by
    the numpy dense product and by scipy, tolerance scaled by sum |a_ik b_kj|."""
    m, k, n = shape
    A = from_sp(sp.random(m, k, density=density, random_state=seed, format="csr",
                          data_rvs=lambda s: np.random.default_rng(seed).standard_normal(s)))
    B = from_sp(sp.random(k, n, density=density, random_state=seed + 100, format="csr",
                          data_rvs=lambda s: np.random.default_rng(seed + 1).standard_normal(s)))
    Cptr, Cidx, Cval, Cabs = oracle.spgemm(A, B, with_abs=True)
    pat = pattern_product(A, B)
    assert [len(r) for r in pat] == np.diff(Cptr).tolist()
    assert sum(pat, []) == Cidx.tolist()
    Dd = dense_of(A.ptr, A.idx, A.val, A.shape) @ dense_of(B.ptr, B.idx, B.val, B.shape)
    rows = np.repeat(np.arange(m), np.diff(Cptr))
    assert np.all(np.abs(Cval - Dd[rows, Cidx]) <= 1e-13 * Cabs + 1e-300)
    S = (A_sp := sp.csr_matrix((A.val, A.idx, A.ptr), shape=A.shape)) @ sp.csr_matrix((B.val, B.idx, B.ptr), shape=B.shape)
    S = S.tocsr()
    Sd = S.toarray()
    assert np.all(np.abs(Cval - Sd[rows, Cidx]) <= 1e-13 * Cabs + 1e-300)
    # abs-sum equals the product of absolute values (a second, independent formula)
    Aabs = dense_of(A.ptr, A.idx, np.abs(A.val), A.shape) @ dense_of(B.ptr, B.idx, np.abs(B.val), B.shape)
    np.testing.assert_allclose(Cabs, Aabs[rows, Cidx], rtol=1e-13)
    del A_sp


def test_structural_zero_kept():
    """A cancelling sum (1*1 + 1*(-1) = 0) stays in the structure (the reference's
    |acc| < 1e-4 first-touch test and CUSP's pass 2 would both lose it)."""
    A = M([0, 2], [0, 1], [1.0, 1.0], (1, 2))
    B = M([0, 1, 2], [0, 0], [1.0, -1.0], (2, 1))
    Cptr, Cidx, Cval = oracle.spgemm(A, B)
    assert Cptr.tolist() == [0, 1] and Cidx.tolist() == [0] and Cval.tolist() == [0.0]
    # tiny partial sums (< 1e-4) are not mistaken for "new" entries
    A = M([0, 3], [0, 1, 2], [1e-5, 1e-5, 1e-5], (1, 3))
    B = M([0, 1, 2, 3], [4, 4, 4], [1.0, 1.0, 1.0], (3, 5))
    Cptr, Cidx, Cval = oracle.spgemm(A, B)
    assert Cidx.tolist() == [4] and abs(Cval[0] - 3e-5) < 1e-20


def test_row_products_two_formulas():
    """P_i (calculate_UpperBound) vs numpy segment sums, and sum(P) vs sum_k colcount_A(k) len(B_k)."""
    A = from_sp(sp.random(300, 200, density=0.03, random_state=9, format="csr"))
    B = from_sp(sp.random(200, 150, density=0.05, random_state=10, format="csr"))
    P = oracle.row_products(A.ptr, A.idx, B.ptr)
    lens = np.diff(B.ptr)
    seg = np.array([lens[A.idx[A.ptr[i]:A.ptr[i + 1]]].sum() for i in range(300)])
    assert P.tolist() == seg.tolist()
    colcount = np.bincount(A.idx, minlength=200)
    assert int(P.sum()) == int((colcount * lens).sum())
    # P bounds nnz: max_k len(B_k) <= nnz_i <= min(P_i, ncols)
    Cptr, Cidx, _ = oracle.spgemm(A, B)
    nnz = np.diff(Cptr)
    assert np.all(nnz <= np.minimum(P, 150))
    for i in range(300):
        ks = A.idx[A.ptr[i]:A.ptr[i + 1]]
        if ks.size:
            assert nnz[i] >= lens[ks].max()


def test_blocks_and_criteria_hand_worked():
    """divide_Blocks + calculate_Criteria + analyze_Block on hand-worked numbers.
    Four 4096-row blocks with 10, 20, 30, 100 products/row: avg = 40; squared deviations
    900 + 400 + 100 + 3600 = 5000; /4 = 1250; sqrt = 35.355; 1.28 * 35.355 = 45.25 -> 45;
    criteria = 85. Blocks at 10/20/30 (<= 64) are CPU-preferred (1), 100 > 85 is GPU-only (0)."""
    BR = 4096
    P = np.repeat(np.array([10, 20, 30, 100], np.int64), BR)
    bi = oracle.blocks(P, BR)
    assert bi.tolist() == [10 * BR, 20 * BR, 30 * BR, 100 * BR]
    c = oracle.criteria(bi, BR)
    assert c == 85
    assert [oracle.analyze_block(c, int(b // BR)) for b in bi] == [1, 1, 1, 0]
    # "any" category: above 64 but not above the criteria
    assert oracle.analyze_block(85, 70) == 2 and oracle.analyze_block(85, 64) == 1 and oracle.analyze_block(85, 86) == 0
    # partial last block: 5000 rows with 512-row blocks -> 10 blocks, last has 392 rows
    P2 = np.arange(5000, dtype=np.int64) % 7
    bi2 = oracle.blocks(P2, 512)
    assert len(bi2) == 10 and bi2[-1] == P2[4608:].sum() and bi2.sum() == P2.sum()


def test_backward_vs_dense_and_finite_differences():
    """Masked gradients: dense float64 closed form dA = (W*mask_C) B^T on pattern(A),
    dB = A^T (W*mask_C) on pattern(B); plus central finite differences of the loss
    L = sum(W .* C) (linear in each entry, so FD is exact up to rounding)."""
    rng = np.random.default_rng(5)
    A = from_sp(sp.random(40, 30, density=0.15, random_state=11, format="csr",
                          data_rvs=lambda s: rng.standard_normal(s)))
    B = from_sp(sp.random(30, 35, density=0.15, random_state=12, format="csr",
                          data_rvs=lambda s: rng.standard_normal(s)))
    Cptr, Cidx, Cval = oracle.spgemm(A, B)
    W = rng.standard_normal(Cidx.size)
    dA, dB, miss = oracle.backward(A, B, Cptr, Cidx, W)
    assert miss == 0
    Wd = dense_of(Cptr, Cidx, W, (40, 35))
    Bd = dense_of(B.ptr, B.idx, B.val, B.shape)
    Ad = dense_of(A.ptr, A.idx, A.val, A.shape)
    GA = Wd @ Bd.T
    GB = Ad.T @ Wd
    ra = np.repeat(np.arange(40), np.diff(A.ptr))
    rb = np.repeat(np.arange(30), np.diff(B.ptr))
    np.testing.assert_allclose(dA, GA[ra, A.idx], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(dB, GB[rb, B.idx], rtol=1e-12, atol=1e-12)

    def loss(Av, Bv):
        _, ci, cv = oracle.spgemm(M(A.ptr, A.idx, Av, A.shape), M(B.ptr, B.idx, Bv, B.shape))
        assert ci.tolist() == Cidx.tolist()
        return float((W * cv).sum())

    h = 1e-3
    for p in rng.choice(A.idx.size, 8, replace=False):
        e = np.zeros(A.idx.size)
        e[p] = h
        fd = (loss(A.val + e, B.val) - loss(A.val - e, B.val)) / (2 * h)
        assert abs(fd - dA[p]) <= 1e-9 * max(1.0, abs(fd))
    for p in rng.choice(B.idx.size, 8, replace=False):
        e = np.zeros(B.idx.size)
        e[p] = h
        fd = (loss(A.val, B.val + e) - loss(A.val, B.val - e)) / (2 * h)
        assert abs(fd - dB[p]) <= 1e-9 * max(1.0, abs(fd))


def test_rows_subset_matches_full():
    A = from_sp(sp.random(200, 200, density=0.04, random_state=13, format="csr"))
    Cptr, Cidx, Cval = oracle.spgemm(A, A)
    sel = np.array([0, 5, 17, 199, 100], np.int32)
    Sptr, Sidx, Sval = oracle.spgemm_rows(sel, A)
    for j, i in enumerate(sel):
        assert Sidx[Sptr[j]:Sptr[j + 1]].tolist() == Cidx[Cptr[i]:Cptr[i + 1]].tolist()
        assert Sval[Sptr[j]:Sptr[j + 1]].tolist() == Cval[Cptr[i]:Cptr[i + 1]].tolist()


def test_planted_mistake_is_caught():
    """Plant a mistake in the oracle's output (drop one product's contribution) and check the
    brute-force comparison used above detects it: the pinning has teeth."""
    A = from_sp(sp.random(20, 20, density=0.2, random_state=21, format="csr"))
    Cptr, Cidx, Cval, Cabs = oracle.spgemm(A, A, with_abs=True)
    bad = Cval.copy()
    bad[len(bad) // 2] -= 0.5 * Cabs[len(bad) // 2]
    Dd = dense_of(A.ptr, A.idx, A.val, A.shape)
    Dd = Dd @ Dd
    rows = np.repeat(np.arange(20), np.diff(Cptr))
    assert np.all(np.abs(Cval - Dd[rows, Cidx]) <= 1e-13 * Cabs)
    assert not np.all(np.abs(bad - Dd[rows, Cidx]) <= 1e-13 * Cabs)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference sources absent (oracle/_ref/libdbdref.so)")
@pytest.mark.parametrize("case", ["cage4", "rand_pos", "pattern", "blocks512"])
def test_oracle_vs_reference_cpu_executor(case):
    """The reference's own CPU executor (Z/computeCPU.cpp, compiled unmodified) on inputs where
    its FP32 / |acc| < 1e-4 first-touch semantics are exact or well-conditioned: every product
    >= 1e-4 and positive, so no partial sum falls below the threshold. After sorting each row,
    structure must equal the oracle's and values agree to FP32 accuracy."""
    rng = np.random.default_rng(1)
    br = 4096
    if case == "cage4":
        A = _mm(os.path.join(GOLD, "cage4.mtx"))
    elif case == "rand_pos":
        A = from_sp(sp.random(700, 700, density=0.02, random_state=2, format="csr",
                              data_rvs=lambda s: rng.uniform(0.1, 1.0, s)))
    elif case == "pattern":
        m = sp.random(1500, 1500, density=0.01, random_state=3, format="csr")
        m.data[:] = 1.0
        A = from_sp(m)
    else:
        A = from_sp(sp.random(3000, 3000, density=0.004, random_state=4, format="csr",
                              data_rvs=lambda s: rng.uniform(0.1, 1.0, s)))
        br = 512
    rows = oracle.ref_cpu_spgemm(A, rows_in_block=br)
    Cptr, Cidx, Cval, Cabs = oracle.spgemm(A, A, with_abs=True)
    for i, (ci, cv) in enumerate(rows):
        o = np.argsort(ci, kind="stable")
        assert ci[o].tolist() == Cidx[Cptr[i]:Cptr[i + 1]].tolist(), f"row {i}"
        err = np.abs(cv[o].astype(np.float64) - Cval[Cptr[i]:Cptr[i + 1]])
        assert np.all(err <= 1e-6 * Cabs[Cptr[i]:Cptr[i + 1]] + 1e-30), f"row {i}"


@pytest.mark.skipif(not oracle.ref_available(), reason="reference sources absent (oracle/_ref/libdbdref.so)")
def test_reference_executor_value_based_first_touch_differs():
    """Why the oracle tracks first touch exactly: with a cancelling partial sum
    (1*1 + 1*(-1) + 1*1 into one column) the reference executor sees |acc| < 1e-4 after the
    second product, appends the column again and ends with a duplicate column entry; the
    oracle keeps one entry with value 1."""
    A = M([0, 3], [0, 1, 2], [1.0, 1.0, 1.0], (1, 3))
    B = M([0, 1, 2, 3], [0, 0, 0], [1.0, -1.0, 1.0], (3, 3))
    rows = oracle.ref_cpu_spgemm(A, B)
    assert rows[0][0].tolist() == [0, 0]  # duplicate column from the reference heuristic
    Cptr, Cidx, Cval = oracle.spgemm(A, B)
    assert Cidx.tolist() == [0] and Cval.tolist() == [1.0]
