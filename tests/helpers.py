"""Shared helpers for the parity tests: run the GPU path through the C ABI and compare with
the oracle. Structure (row pointer, column indices) must match exactly; values within
|gpu - ref| <= rtol * sum_k |a_ik b_kj| (the error scale of any summation order), with
rtol = 1e-12 for FP64 and 1e-5 for FP32."""
from __future__ import annotations

import numpy as np

import dbd_amd
import oracle

RTOL = {np.dtype(np.float64): 1e-12, np.dtype(np.float32): 1e-5}


def reference(A, B=None):
    Cptr, Cidx, Cval, Cabs = oracle.spgemm(A, B if B is not None else A, with_abs=True)
    return Cptr, Cidx, Cval, Cabs


def compare(C, ref, dtype, what=""):
    Cptr, Cidx, Cval, Cabs = ref
    assert C.ptr.shape == Cptr.shape, f"{what}: row count"
    bad_rows = np.nonzero(C.ptr != Cptr)[0]
    assert bad_rows.size == 0, f"{what}: row pointer differs first at row {bad_rows[:5]}"
    mism = np.nonzero(C.idx != Cidx)[0]
    assert mism.size == 0, f"{what}: {mism.size} column indices differ, first at {mism[:5]}"
    rtol = RTOL[np.dtype(dtype)]
    err = np.abs(C.val.astype(np.float64) - Cval)
    lim = rtol * Cabs + 1e-300
    badv = np.nonzero(err > lim)[0]
    assert badv.size == 0, (f"{what}: {badv.size} values off, worst rel "
                            f"{float(np.max(err / np.maximum(Cabs, 1e-300))):.3e}")


def gpu_spgemm(A, B=None, mode="compute", block_rows=4096, table=None, **kw):
    """Full pipeline run (with candidate table `table` when given); returns (host CSR, timing,
    plan info)."""
    with dbd_amd.Context(block_rows=block_rows, dtype=A.val.dtype, **kw) as ctx:
        if table is not None:
            ctx.set_table(table)
        ctx.set_operands(A, B)
        t = ctx.run(mode)
        info = ctx.plan_info()
        C = ctx.result() if mode != "io" else None
        return C, t, info


def check_product(A, B=None, what="", **kw):
    ref = reference(A, B)
    C, t, info = gpu_spgemm(A, B, **kw)
    compare(C, ref, A.val.dtype, what)
    return C, t, info
