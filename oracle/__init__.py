"""ctypes binding of the CPU oracle (oracle.c). TEST INFRASTRUCTURE ONLY: imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the dbd_amd product.

All functions take numpy arrays (int64 row pointers, int32 columns, float64 values) and
return numpy results. Build with `make -C oracle` (outputs into oracle/_ref/).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libdbdref.so")


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB):
        build()
    lb = C.CDLL(LIB)
    P, I64, I32 = C.c_void_p, C.c_int64, C.c_int32
    lb.oracle_row_products.argtypes = [I64, P, P, P, P]
    lb.oracle_row_products.restype = None
    lb.oracle_blocks.argtypes = [P, I64, I32, P]
    lb.oracle_blocks.restype = I64
    lb.oracle_criteria.argtypes = [P, I64, I32]
    lb.oracle_criteria.restype = I64
    lb.oracle_analyze_block.argtypes = [I64, I64]
    lb.oracle_analyze_block.restype = I32
    lb.oracle_spgemm_count.argtypes = [I64, I64, P, P, P, P, P, C.c_int]
    lb.oracle_spgemm_count.restype = I64
    lb.oracle_spgemm_fill.argtypes = [I64, I64, P, P, P, P, P, P, P, P, P, P, C.c_int]
    lb.oracle_spgemm_fill.restype = None
    lb.oracle_spgemm_rows.argtypes = [I64, P, I64, P, P, P, P, P, P, P, P, P, P, C.c_int]
    lb.oracle_spgemm_rows.restype = I64
    lb.oracle_backward.argtypes = [I64, P, P, P, P, P, P, P, P, P, P, P, I64]
    lb.oracle_backward.restype = I64
    return lb


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def row_products(Aptr, Aidx, Bptr) -> np.ndarray:
    Aptr, Aidx, Bptr = _c(Aptr, np.int64), _c(Aidx, np.int32), _c(Bptr, np.int64)
    nr = Aptr.shape[0] - 1
    P = np.zeros(nr, np.int64)
    lib().oracle_row_products(nr, Aptr.ctypes.data, Aidx.ctypes.data, Bptr.ctypes.data, P.ctypes.data)
    return P


def blocks(P, block_rows: int) -> np.ndarray:
    P = _c(P, np.int64)
    nr = P.shape[0]
    out = np.zeros(max(1, (nr - 1) // block_rows + 1 if nr else 1), np.int64)
    nb = lib().oracle_blocks(P.ctypes.data, nr, block_rows, out.ctypes.data)
    return out[:nb]


def criteria(block_info, block_rows: int) -> int:
    bi = _c(block_info, np.int64)
    return int(lib().oracle_criteria(bi.ctypes.data, bi.shape[0], block_rows))


def analyze_block(criteria_: int, avg: int) -> int:
    return int(lib().oracle_analyze_block(criteria_, avg))


def spgemm(A, B=None, threads: int = 0, with_abs: bool = False):
    """C = A*B in FP64. A, B: objects with ptr / idx / val / shape (e.g. dbd_amd.CSR) or
    (ptr, idx, val, shape) tuples. Returns (Cptr, Cidx, Cval[, Cabs])."""
    def unpack(M):
        if isinstance(M, tuple):
            return M
        return M.ptr, M.idx, M.val, M.shape
    Aptr, Aidx, Aval, Ash = unpack(A)
    Bptr, Bidx, Bval, Bsh = unpack(B if B is not None else A)
    assert Ash[1] == Bsh[0], "inner dimensions differ"
    Aptr, Aidx, Aval = _c(Aptr, np.int64), _c(Aidx, np.int32), _c(Aval, np.float64)
    Bptr, Bidx, Bval = _c(Bptr, np.int64), _c(Bidx, np.int32), _c(Bval, np.float64)
    nr, nc = int(Ash[0]), int(Bsh[1])
    Cptr = np.zeros(nr + 1, np.int64)
    L = lib()
    nnz = L.oracle_spgemm_count(nr, nc, Aptr.ctypes.data, Aidx.ctypes.data, Bptr.ctypes.data, Bidx.ctypes.data,
                                Cptr.ctypes.data, threads)
    Cidx = np.zeros(max(nnz, 1), np.int32)
    Cval = np.zeros(max(nnz, 1), np.float64)
    Cabs = np.zeros(max(nnz, 1), np.float64) if with_abs else None
    L.oracle_spgemm_fill(nr, nc, Aptr.ctypes.data, Aidx.ctypes.data, Aval.ctypes.data, Bptr.ctypes.data,
                         Bidx.ctypes.data, Bval.ctypes.data, Cptr.ctypes.data, Cidx.ctypes.data, Cval.ctypes.data,
                         Cabs.ctypes.data if with_abs else None, threads)
    if with_abs:
        return Cptr, Cidx[:nnz], Cval[:nnz], Cabs[:nnz]
    return Cptr, Cidx[:nnz], Cval[:nnz]


def spgemm_rows(rows, A, B=None, threads: int = 0, with_abs: bool = False):
    """Rows `rows` of A*B only: returns (Sptr, Sidx, Sval[, Sabs]) over the selected rows."""
    Bm = B if B is not None else A
    sel = _c(rows, np.int32)
    Aptr, Aidx, Aval = _c(A.ptr, np.int64), _c(A.idx, np.int32), _c(A.val, np.float64)
    Bptr, Bidx, Bval = _c(Bm.ptr, np.int64), _c(Bm.idx, np.int32), _c(Bm.val, np.float64)
    Sptr = np.zeros(sel.shape[0] + 1, np.int64)
    L = lib()
    nnz = L.oracle_spgemm_rows(sel.shape[0], sel.ctypes.data, Bm.shape[1], Aptr.ctypes.data, Aidx.ctypes.data,
                               Aval.ctypes.data, Bptr.ctypes.data, Bidx.ctypes.data, Bval.ctypes.data,
                               Sptr.ctypes.data, None, None, None, threads)
    Sidx = np.zeros(max(nnz, 1), np.int32)
    Sval = np.zeros(max(nnz, 1), np.float64)
    Sabs = np.zeros(max(nnz, 1), np.float64) if with_abs else None
    L.oracle_spgemm_rows(sel.shape[0], sel.ctypes.data, Bm.shape[1], Aptr.ctypes.data, Aidx.ctypes.data,
                         Aval.ctypes.data, Bptr.ctypes.data, Bidx.ctypes.data, Bval.ctypes.data, Sptr.ctypes.data,
                         Sidx.ctypes.data, Sval.ctypes.data, Sabs.ctypes.data if with_abs else None, threads)
    if with_abs:
        return Sptr, Sidx[:nnz], Sval[:nnz], Sabs[:nnz]
    return Sptr, Sidx[:nnz], Sval[:nnz]


def backward(A, B, Cptr, Cidx, dC):
    """FP64 masked gradients: returns (dA values on pattern(A), dB values on pattern(B), missing)."""
    Aptr, Aidx, Aval = _c(A.ptr, np.int64), _c(A.idx, np.int32), _c(A.val, np.float64)
    Bptr, Bidx, Bval = _c(B.ptr, np.int64), _c(B.idx, np.int32), _c(B.val, np.float64)
    Cptr, Cidx, dC = _c(Cptr, np.int64), _c(Cidx, np.int32), _c(dC, np.float64)
    dA = np.zeros(max(Aidx.shape[0], 1), np.float64)
    dB = np.zeros(max(Bidx.shape[0], 1), np.float64)
    miss = lib().oracle_backward(Aptr.shape[0] - 1, Aptr.ctypes.data, Aidx.ctypes.data, Aval.ctypes.data,
                                 Bptr.ctypes.data, Bidx.ctypes.data, Bval.ctypes.data, Cptr.ctypes.data,
                                 Cidx.ctypes.data, dC.ctypes.data, dA.ctypes.data, dB.ctypes.data, Bidx.shape[0])
    return dA[: Aidx.shape[0]], dB[: Bidx.shape[0]], int(miss)


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref_cpu_spgemm(A, B=None, rows_in_block: int = 4096):
    """C = A*B by the reference's own CPU executor (compute_CPU of Z/computeCPU.cpp, FP32,
    |acc| < 1e-4 first touch, first-touch column order), built by `make -C oracle ref`.
    Returns per-row lists (cols, vals) in the executor's order."""
    Bm = B if B is not None else A
    lb = C.CDLL(REF_LIB)
    P_ = C.c_void_p
    lb.dbdref_spgemm.argtypes = [C.c_int, C.c_int, C.c_int, P_, P_, P_, P_, P_, P_, C.c_int, P_, P_, P_, P_]
    lb.dbdref_spgemm.restype = C.c_int
    nr = int(A.shape[0])
    Ap, Ai, Av = _c(A.ptr, np.int32), _c(A.idx, np.int32), _c(A.val, np.float32)
    Bp, Bi, Bv = _c(Bm.ptr, np.int32), _c(Bm.idx, np.int32), _c(Bm.val, np.float32)
    P = row_products(A.ptr, A.idx, Bm.ptr)
    off = np.zeros(nr + 1, np.int64)
    off[1:] = np.cumsum(P)
    cnt = np.zeros(nr, np.int32)
    oi = np.zeros(max(int(off[-1]), 1), np.int32)
    ov = np.zeros(max(int(off[-1]), 1), np.float32)
    rc = lb.dbdref_spgemm(nr, int(A.shape[1]), int(Bm.shape[1]), Ap.ctypes.data, Ai.ctypes.data, Av.ctypes.data,
                          Bp.ctypes.data, Bi.ctypes.data, Bv.ctypes.data, rows_in_block, cnt.ctypes.data,
                          off.ctypes.data, oi.ctypes.data, ov.ctypes.data)
    if rc != 0:
        raise RuntimeError("reference executor failed")
    return [(oi[off[r]:off[r] + cnt[r]].copy(), ov[off[r]:off[r] + cnt[r]].copy()) for r in range(nr)]
