"""ctypes declarations mirroring include/dbd.h exactly (struct layouts and signatures)."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# DBD_LIBRARY selects another build of the same library (an A/B variant built with
# `python dbd_amd/build.py --lib-out PATH -D NAME=VALUE`); default: the in-tree build.
lib_path = os.environ.get("DBD_LIBRARY") or os.path.join(HERE, "libdbd.so")

MODE_COMPUTE, MODE_D2H, MODE_IO = 0, 1, 2
F32, F64 = 0, 1
HOST, DEVICE = 0, 1
MAX_CATS = 16

STATUS_NAMES = {0: "DBD_OK", 1: "DBD_ERR_INVALID", 2: "DBD_ERR_HIP", 3: "DBD_ERR_NOMEM", 4: "DBD_ERR_BUDGET",
                5: "DBD_ERR_IO", 6: "DBD_ERR_OVERFLOW", 7: "DBD_ERR_STATE", 8: "DBD_ERR_INTERNAL"}


class DBDError(RuntimeError):
    def __init__(self, status: int, msg: str):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")


class dbd_csr(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("ncols", C.c_int64), ("nnz", C.c_int64),
                ("ptr", C.c_void_p), ("idx", C.c_void_p), ("val", C.c_void_p),
                ("dtype", C.c_int32), ("location", C.c_int32)]


class dbd_opts(C.Structure):
    _fields_ = [("block_rows", C.c_int32), ("dtype", C.c_int32), ("host_budget_mib", C.c_int64),
                ("device_budget_mib", C.c_int64), ("tune", C.c_int32), ("verbose", C.c_int32),
                ("out_dir", C.c_char_p), ("ptr64_file", C.c_int32), ("writer_threads", C.c_int32),
                ("chunks_hint", C.c_int32), ("retain_result", C.c_int32), ("hybrid", C.c_int32),
                ("cpu_threads", C.c_int32)]


class dbd_plan_info(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("ncols", C.c_int64), ("num_blocks", C.c_int64), ("block_rows", C.c_int64),
                ("products", C.c_int64), ("max_block_products", C.c_int64), ("criteria", C.c_int64),
                ("blocks_gpu_only", C.c_int64), ("blocks_cpu_pref", C.c_int64), ("blocks_any", C.c_int64),
                ("nnz", C.c_int64), ("n_sym_cats", C.c_int32), ("n_num_cats", C.c_int32),
                ("sym_rows", C.c_int64 * MAX_CATS), ("num_rows", C.c_int64 * MAX_CATS),
                ("sym_limit", C.c_int64 * MAX_CATS), ("num_limit", C.c_int64 * MAX_CATS)]

    def to_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if not k.endswith(("_rows", "_limit")) or k == "block_rows"}
        d["sym_rows"] = list(self.sym_rows)[: self.n_sym_cats]
        d["num_rows"] = list(self.num_rows)[: self.n_num_cats]
        d["sym_limit"] = list(self.sym_limit)[: self.n_sym_cats]
        d["num_limit"] = list(self.num_limit)[: self.n_num_cats]
        return d


class dbd_table_desc(C.Structure):
    _fields_ = [("g_small", C.c_int32), ("g_big", C.c_int32), ("load", C.c_double), ("top", C.c_int32),
                ("hub_min_cols", C.c_int64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class dbd_timing(C.Structure):
    _fields_ = [("preprocess_s", C.c_double), ("compute_gpu_s", C.c_double), ("write_s", C.c_double),
                ("kernel_exec_s", C.c_double), ("total_s", C.c_double), ("h2d_s", C.c_double),
                ("d2h_s", C.c_double), ("d2h_bytes", C.c_double), ("io_bytes", C.c_double),
                ("chunks", C.c_int64), ("symbolic_s", C.c_double), ("numeric_s", C.c_double),
                ("cpu_s", C.c_double), ("cpu_blocks", C.c_int64), ("cpu_products", C.c_int64)]

    def to_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


P = C.c_void_p
I32, I64, U64, DBL = C.c_int32, C.c_int64, C.c_uint64, C.c_double
CSRP = C.POINTER(dbd_csr)

# name -> (restype, argtypes)
SIGNATURES = {
    "dbd_abi_version": (I32, []),
    "dbd_last_error": (C.c_char_p, []),
    "dbd_default_opts": (None, [C.POINTER(dbd_opts)]),
    "dbd_create": (I32, [C.POINTER(dbd_opts), C.POINTER(P)]),
    "dbd_destroy": (I32, [P]),
    "dbd_set_stream": (I32, [P, P]),
    "dbd_use_own_stream": (I32, [P]),
    "dbd_set_operands": (I32, [P, CSRP, CSRP]),
    "dbd_plan": (I32, [P, C.POINTER(dbd_plan_info)]),
    "dbd_run": (I32, [P, I32, C.POINTER(dbd_timing)]),
    "dbd_result": (I32, [P, CSRP]),
    "dbd_copy_result": (I32, [P, P, P, P]),
    "dbd_get_plan_info": (I32, [P, C.POINTER(dbd_plan_info)]),
    "dbd_get_row_products": (I32, [P, P]),
    "dbd_get_blocks": (I32, [P, P, P]),
    "dbd_symbolic": (I32, [P, P, C.POINTER(I64)]),
    "dbd_numeric": (I32, [P, P, P, P]),
    "dbd_backward": (I32, [P, P, P, P, P, P]),
    "dbd_batch_prepare": (I32, [P, CSRP, CSRP, C.POINTER(I64)]),
    "dbd_batch_launch": (I32, [P, P, P, P]),
    "dbd_autotune": (I32, [P, C.POINTER(I32)]),
    "dbd_table_count": (I32, []),
    "dbd_table_knobs": (I32, [I32, C.POINTER(dbd_table_desc)]),
    "dbd_set_table": (I32, [P, I32]),
    "dbd_noalloc_scratch_bytes": (C.c_size_t, [I64, I64]),
    "dbd_spgemm_noalloc": (I32, [CSRP, CSRP, P, P, P, P, I64, C.POINTER(I64)]),
    "dbd_plan_blocks": (I32, [P, I64, I32, P, P, C.POINTER(dbd_plan_info)]),
    "dbd_analyze_block": (I32, [I64, I64]),
    "dbd_read_matrix": (I32, [C.c_char_p, I32, I32, CSRP]),
    "dbd_write_csr_cache": (I32, [C.c_char_p, CSRP]),
    "dbd_csr_free": (I32, [CSRP]),
    "dbd_write_outputs": (I32, [C.c_char_p, CSRP, I32]),
    "dbd_gen_er": (I32, [I64, I64, I32, U64, I32, CSRP]),
    "dbd_gen_web": (I32, [I64, DBL, I32, I32, I64, DBL, U64, I32, CSRP]),
    "dbd_gen_rmat": (I32, [I32, I32, DBL, DBL, DBL, U64, I32, I32, I32, CSRP]),
    "dbd_gen_skewed": (I32, [I64, DBL, DBL, U64, I32, CSRP]),
}


def _load() -> C.CDLL:
    if not os.path.exists(lib_path):
        raise ImportError(f"{lib_path} is not built; run `python dbd_amd/build.py` (hipcc, gfx950)")
    # One HIP runtime per process: torch ships its own libamdhip64 and asks for it by the
    # unversioned name, so if libdbd loaded the system copy first, torch would load a second
    # runtime and device pointers would not be shared. Loading torch first makes libdbd's
    # libamdhip64.so.7 resolve to torch's copy (same soname).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lb = C.CDLL(lib_path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lb, name)
        f.restype = res
        f.argtypes = args
    if lb.dbd_abi_version() != 3:
        raise ImportError("libdbd ABI version mismatch")
    return lb


lib = _load()


def check(status: int) -> None:
    if status != 0:
        raise DBDError(status, lib.dbd_last_error().decode(errors="replace"))
