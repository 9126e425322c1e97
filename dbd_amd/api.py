"""Python API over the C ABI: host CSR container, generators, file formats, the pipeline
context (the reference's compute / d2h / SSpMM modes) and the device two-phase API used by the
torch binding. Device memory is plumbed through torch tensors; all compute is in libdbd.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from ._lib import (DEVICE, F32, F64, HOST, MODE_COMPUTE, MODE_D2H, MODE_IO, check, dbd_csr, dbd_opts, dbd_plan_info,
                   dbd_table_desc, dbd_timing, lib)

_MODES = {"compute": MODE_COMPUTE, "d2h": MODE_D2H, "io": MODE_IO, "sspmm": MODE_IO,
          MODE_COMPUTE: MODE_COMPUTE, MODE_D2H: MODE_D2H, MODE_IO: MODE_IO}


def _dtype_code(dtype) -> int:
    if dtype in (F32, F64):
        return dtype
    d = np.dtype(dtype) if not isinstance(dtype, str) or dtype not in ("f32", "f64") else np.dtype(
        {"f32": np.float32, "f64": np.float64}[dtype])
    if d == np.float64:
        return F64
    if d == np.float32:
        return F32
    raise ValueError(f"unsupported value dtype {dtype}")


def _np_dtype(code: int):
    return np.float64 if code == F64 else np.float32


@dataclass
class CSR:
    """Host CSR: ptr int64[nrows+1], idx int32[nnz] (strictly increasing per row), val f32/f64."""
    ptr: np.ndarray
    idx: np.ndarray
    val: np.ndarray
    shape: tuple

    def __post_init__(self):
        self.ptr = np.ascontiguousarray(self.ptr, dtype=np.int64)
        self.idx = np.ascontiguousarray(self.idx, dtype=np.int32)
        if self.val.dtype not in (np.float32, np.float64):
            self.val = self.val.astype(np.float64)
        self.val = np.ascontiguousarray(self.val)
        self.shape = (int(self.shape[0]), int(self.shape[1]))

    @property
    def nnz(self) -> int:
        return int(self.ptr[-1])

    @property
    def dtype_code(self) -> int:
        return _dtype_code(self.val.dtype)

    def astype(self, dtype) -> "CSR":
        return CSR(self.ptr, self.idx, self.val.astype(_np_dtype(_dtype_code(dtype))), self.shape)

    def view(self) -> dbd_csr:
        v = dbd_csr()
        v.nrows, v.ncols, v.nnz = self.shape[0], self.shape[1], self.nnz
        v.ptr = self.ptr.ctypes.data
        v.idx = self.idx.ctypes.data
        v.val = self.val.ctypes.data
        v.dtype = self.dtype_code
        v.location = HOST
        return v

    @classmethod
    def from_scipy(cls, m, dtype=np.float64) -> "CSR":
        m = m.tocsr()
        m.sort_indices()
        m.sum_duplicates()
        return cls(m.indptr.astype(np.int64), m.indices.astype(np.int32), m.data.astype(dtype), m.shape)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.idx, self.ptr), shape=self.shape)

    def to_dense(self) -> np.ndarray:
        d = np.zeros(self.shape, dtype=self.val.dtype)
        rows = np.repeat(np.arange(self.shape[0]), np.diff(self.ptr))
        np.add.at(d, (rows, self.idx), self.val)
        return d


@dataclass
class DeviceCSR:
    """CSR whose arrays live in device memory (torch tensors on the GPU); borrowed by libdbd."""
    ptr: object  # torch.int64 [nrows+1]
    idx: object  # torch.int32 [nnz]
    val: object  # torch.float32 / float64 [nnz]
    shape: tuple

    @property
    def nnz(self) -> int:
        return int(self.idx.numel())

    def view(self) -> dbd_csr:
        import torch
        assert self.ptr.dtype == torch.int64 and self.idx.dtype == torch.int32
        assert self.ptr.is_cuda and self.idx.is_cuda and self.val.is_cuda
        v = dbd_csr()
        v.nrows, v.ncols, v.nnz = int(self.shape[0]), int(self.shape[1]), self.nnz
        v.ptr, v.idx, v.val = self.ptr.data_ptr(), self.idx.data_ptr(), self.val.data_ptr()
        v.dtype = F64 if self.val.dtype == torch.float64 else F32
        v.location = DEVICE
        return v

    @classmethod
    def from_host(cls, M: CSR, device="cuda") -> "DeviceCSR":
        import torch
        return cls(torch.from_numpy(M.ptr).to(device), torch.from_numpy(M.idx).to(device),
                   torch.from_numpy(M.val).to(device), M.shape)


def _take_csr(v: dbd_csr) -> CSR:
    """Copy a library-allocated host CSR into numpy and free it."""
    try:
        nr, nnz = v.nrows, v.nnz
        ptr = np.ctypeslib.as_array(C.cast(v.ptr, C.POINTER(C.c_int64)), shape=(nr + 1,)).copy()
        idx = (np.ctypeslib.as_array(C.cast(v.idx, C.POINTER(C.c_int32)), shape=(nnz,)).copy() if nnz
               else np.zeros(0, np.int32))
        ct = C.c_double if v.dtype == F64 else C.c_float
        val = (np.ctypeslib.as_array(C.cast(v.val, C.POINTER(ct)), shape=(nnz,)).copy() if nnz
               else np.zeros(0, _np_dtype(v.dtype)))
        return CSR(ptr, idx, val, (nr, v.ncols))
    finally:
        lib.dbd_csr_free(C.byref(v))


# ---- generators (seeded, deterministic) ------------------------------------------------
def gen_er(nrows: int, ncols: int, deg: int, seed: int = 1, dtype=np.float64) -> CSR:
    v = dbd_csr()
    check(lib.dbd_gen_er(nrows, ncols, deg, seed, _dtype_code(dtype), C.byref(v)))
    return _take_csr(v)


def gen_web(n: int, zipf_s: float = 2.1, dmin: int = 1, dmax: int = 255, band: int = 2000, p_local: float = 0.7,
            seed: int = 1, dtype=np.float64) -> CSR:
    v = dbd_csr()
    check(lib.dbd_gen_web(n, zipf_s, dmin, dmax, band, p_local, seed, _dtype_code(dtype), C.byref(v)))
    return _take_csr(v)


def gen_rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19, seed: int = 1,
             add_identity: bool = False, row_normalize: bool = False, dtype=np.float64) -> CSR:
    v = dbd_csr()
    check(lib.dbd_gen_rmat(scale, edge_factor, a, b, c, seed, int(add_identity), int(row_normalize),
                           _dtype_code(dtype), C.byref(v)))
    return _take_csr(v)


def gen_skewed(n: int, mean_deg: float, zipf_s: float = 1.8, seed: int = 1, dtype=np.float32) -> CSR:
    v = dbd_csr()
    check(lib.dbd_gen_skewed(n, mean_deg, zipf_s, seed, _dtype_code(dtype), C.byref(v)))
    return _take_csr(v)


# ---- host-side file formats + plan logic ------------------------------------------------
def read_matrix(path: str, dtype=np.float32, use_cache: bool = False) -> CSR:
    v = dbd_csr()
    check(lib.dbd_read_matrix(path.encode(), _dtype_code(dtype), int(use_cache), C.byref(v)))
    return _take_csr(v)


def write_csr_cache(path: str, M: CSR) -> None:
    v = M.view()
    check(lib.dbd_write_csr_cache(path.encode(), C.byref(v)))


def write_outputs(directory: str, M: CSR, ptr64: bool = False) -> None:
    v = M.view()
    check(lib.dbd_write_outputs(directory.encode(), C.byref(v), int(ptr64)))


def plan_blocks(row_products: np.ndarray, block_rows: int):
    P = np.ascontiguousarray(row_products, dtype=np.int64)
    nr = P.shape[0]
    nb = 0 if nr == 0 else (nr - 1) // block_rows + 1
    bptr = np.zeros(nb + 1, np.int64)
    bprod = np.zeros(max(nb, 1), np.int64)
    info = dbd_plan_info()
    check(lib.dbd_plan_blocks(P.ctypes.data, nr, block_rows, bptr.ctypes.data, bprod.ctypes.data, C.byref(info)))
    return bptr, bprod[:nb], info.to_dict()


def analyze_block(criteria: int, avg_products_per_row: int) -> int:
    return int(lib.dbd_analyze_block(criteria, avg_products_per_row))


# ---- the pipeline -----------------------------------------------------------------------
class Context:
    """One dbd_ctx: operands, plan, runs. Mirrors the reference's binaries' knobs: block rows
    (ROWS_IN_BLOCK), host and device DRAM budgets in GiB, and the three modes. hybrid=True runs
    the blocks the reference's analyze_Block sends to the CPU on host threads (D2H and IO modes;
    the operands must be host CSR)."""

    def __init__(self, block_rows: int = 4096, dtype=np.float64, host_gib: float = 28, device_gib: float = 22,
                 tune: bool = False, out_dir: Optional[str] = None, ptr64: bool = False, writer_threads: int = 4,
                 chunks_hint: int = 8, retain: bool = True, stream=None, hybrid: bool = False,
                 cpu_threads: int = 0):
        o = dbd_opts()
        lib.dbd_default_opts(C.byref(o))
        o.block_rows = block_rows
        o.dtype = _dtype_code(dtype)
        o.host_budget_mib = int(host_gib * 1024)
        o.device_budget_mib = int(device_gib * 1024)
        o.tune = int(tune)
        self._out_dir = out_dir.encode() if out_dir else None
        o.out_dir = self._out_dir
        o.ptr64_file = int(ptr64)
        o.writer_threads = writer_threads
        o.chunks_hint = chunks_hint
        o.retain_result = int(retain)
        o.hybrid = int(hybrid)
        o.cpu_threads = int(cpu_threads)
        self.dtype = o.dtype
        h = C.c_void_p()
        check(lib.dbd_create(C.byref(o), C.byref(h)))
        self._h = h
        self._keep = []
        if stream is not None:
            self.set_stream(stream)

    def close(self):
        if getattr(self, "_h", None):
            lib.dbd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream) -> None:
        """stream: a torch.cuda.Stream or a raw hipStream_t int (0 is the default stream, used as
        such, so the work stays ordered with torch's default-stream work), or None (a stream the
        context owns)."""
        if stream is None:
            check(lib.dbd_use_own_stream(self._h))
            return
        check(lib.dbd_set_stream(self._h, int(getattr(stream, "cuda_stream", stream))))

    def set_operands(self, A, B=None) -> None:
        va = A.view()
        vb = va if B is None or B is A else B.view()
        self._keep = [A, B, va, vb]
        check(lib.dbd_set_operands(self._h, C.byref(va), C.byref(vb)))

    def plan(self) -> dict:
        info = dbd_plan_info()
        check(lib.dbd_plan(self._h, C.byref(info)))
        return info.to_dict()

    def plan_info(self) -> dict:
        info = dbd_plan_info()
        check(lib.dbd_get_plan_info(self._h, C.byref(info)))
        return info.to_dict()

    def row_products(self, nrows: int) -> np.ndarray:
        P = np.zeros(nrows, np.int64)
        check(lib.dbd_get_row_products(self._h, P.ctypes.data))
        return P

    def blocks(self, num_blocks: int):
        bptr = np.zeros(num_blocks + 1, np.int64)
        bprod = np.zeros(max(num_blocks, 1), np.int64)
        check(lib.dbd_get_blocks(self._h, bptr.ctypes.data, bprod.ctypes.data))
        return bptr, bprod[:num_blocks]

    def run(self, mode="compute") -> dict:
        t = dbd_timing()
        check(lib.dbd_run(self._h, _MODES[mode], C.byref(t)))
        return t.to_dict()

    def result_view(self) -> dbd_csr:
        v = dbd_csr()
        check(lib.dbd_result(self._h, C.byref(v)))
        return v

    def result(self) -> CSR:
        """Copy of the retained result (device or pinned host) as a host CSR; the library does
        the copy on its own stream."""
        v = self.result_view()
        nr, nnz = v.nrows, v.nnz
        ptr = np.empty(nr + 1, np.int64)
        idx = np.empty(nnz, np.int32)
        val = np.empty(nnz, _np_dtype(v.dtype))
        check(lib.dbd_copy_result(self._h, ptr.ctypes.data, idx.ctypes.data if nnz else None,
                                  val.ctypes.data if nnz else None))
        return CSR(ptr, idx, val, (nr, v.ncols))

    # -- device two-phase API (torch tensors) --
    def symbolic(self, Cptr) -> int:
        n = C.c_int64()
        check(lib.dbd_symbolic(self._h, Cptr.data_ptr() if Cptr is not None else None, C.byref(n)))
        return int(n.value)

    def numeric(self, Cptr, Cidx, Cval) -> None:
        check(lib.dbd_numeric(self._h, Cptr.data_ptr(), Cidx.data_ptr() if Cidx.numel() else None,
                              Cval.data_ptr() if Cval.numel() else None))

    def backward(self, Cptr, Cidx, dC, dA=None, dB=None) -> None:
        check(lib.dbd_backward(self._h, Cptr.data_ptr(), Cidx.data_ptr(), dC.data_ptr(),
                               dA.data_ptr() if dA is not None else None, dB.data_ptr() if dB is not None else None))

    @staticmethod
    def table_count() -> int:
        return int(lib.dbd_table_count())

    @staticmethod
    def table_knobs(table_id: int) -> dict:
        """Knobs of candidate table `table_id` (table 0 is the default): lanes per B row in small /
        large hash tables, the load limit, the largest numeric LDS table, and the column count from
        which hub rows take the numeric hub kernels."""
        d = dbd_table_desc()
        check(lib.dbd_table_knobs(table_id, C.byref(d)))
        return d.as_dict()

    def set_table(self, table_id: int) -> None:
        check(lib.dbd_set_table(self._h, table_id))

    def autotune(self) -> int:
        ch = C.c_int32()
        check(lib.dbd_autotune(self._h, C.byref(ch)))
        return int(ch.value)

    def batch_prepare(self, A_cat, B_cat) -> int:
        """Plan the batch and run its symbolic pass once; returns the exact nnz of C_cat."""
        va, vb = A_cat.view(), B_cat.view()
        self._keep = [A_cat, B_cat, va, vb]
        n = C.c_int64()
        check(lib.dbd_batch_prepare(self._h, C.byref(va), C.byref(vb), C.byref(n)))
        return int(n.value)

    def batch_launch(self, Cptr, Cidx, Cval) -> None:
        check(lib.dbd_batch_launch(self._h, Cptr.data_ptr(), Cidx.data_ptr(), Cval.data_ptr()))


def spgemm(A: CSR, B: Optional[CSR] = None, mode="d2h", block_rows: int = 4096, **kw):
    """C = A*B through the full pipeline; returns (C as host CSR or None, timing dict).
    mode "compute" keeps C on the device (copied back here), "d2h" into pinned host memory,
    "io" writes Ptr/Idx/Val.output (returns None for C)."""
    dtype = kw.pop("dtype", A.val.dtype)
    with Context(block_rows=block_rows, dtype=dtype, **kw) as ctx:
        ctx.set_operands(A, B)
        t = ctx.run(mode)
        Cm = None if _MODES[mode] == MODE_IO else ctx.result()
        return Cm, t
