"""The entry point that allocates nothing (dbd_spgemm_noalloc), called through the C ABI, against
the batched form's existing entry points (dbd_batch_prepare + dbd_batch_launch, the same pipeline
with the context's own workspace) and the oracle: structure bit-exact, values within
rtol * sum |a_ik b_kj| (rtol 1e-12 FP64, 1e-5 FP32). Two calls in a row, the second on a
smaller input, so the fixed workspace holds the first call's contents when the second starts.
Also: a capacity that is too small and a problem too large for the fixed workspace are refused
(on the host, before any device work, for the latter)."""
import ctypes as C

import numpy as np
import pytest

import dbd_amd
from dbd_amd import _lib
from helpers import compare, reference

pytestmark = pytest.mark.gpu


def run_noalloc(A, B, capacity=None):
    import torch
    dA = dbd_amd.DeviceCSR.from_host(A)
    dB = dA if B is None else dbd_amd.DeviceCSR.from_host(B)
    va, vb = dA.view(), dB.view()
    tdt = torch.float64 if A.val.dtype == np.float64 else torch.float32
    P = int(np.diff((B if B is not None else A).ptr)[A.idx].sum())
    cap = P if capacity is None else capacity
    Cptr = torch.empty(A.shape[0] + 1, dtype=torch.int64, device="cuda")
    Cidx = torch.empty(max(cap, 1), dtype=torch.int32, device="cuda")
    Cval = torch.empty(max(cap, 1), dtype=tdt, device="cuda")
    s = torch.cuda.current_stream()
    nnz = C.c_int64(-1)
    st = _lib.lib.dbd_spgemm_noalloc(C.byref(va), C.byref(vb), s.cuda_stream, Cptr.data_ptr(), Cidx.data_ptr(),
                                     Cval.data_ptr(), cap, C.byref(nnz))
    return st, int(nnz.value), Cptr, Cidx, Cval


def run_batch_entry(A, B):
    import torch
    dA = dbd_amd.DeviceCSR.from_host(A)
    dB = dA if B is None else dbd_amd.DeviceCSR.from_host(B)
    tdt = torch.float64 if A.val.dtype == np.float64 else torch.float32
    with dbd_amd.Context(block_rows=512, dtype=A.val.dtype) as ctx:
        nnz = ctx.batch_prepare(dA, dB)
        Cptr = torch.empty(A.shape[0] + 1, dtype=torch.int64, device="cuda")
        Cidx = torch.empty(max(nnz, 1), dtype=torch.int32, device="cuda")
        Cval = torch.empty(max(nnz, 1), dtype=tdt, device="cuda")
        ctx.batch_launch(Cptr, Cidx, Cval)
        torch.cuda.synchronize()
        return dbd_amd.CSR(Cptr.cpu().numpy(), Cidx[:nnz].cpu().numpy(), Cval[:nnz].cpu().numpy(), A.shape)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_matches_existing_entry_point_twice(gpu_ok, dtype):
    first = (dbd_amd.gen_rmat(14, 16, seed=41, dtype=dtype), None)  # every category incl. long rows
    second = (dbd_amd.gen_er(3000, 2500, 12, seed=42, dtype=dtype), dbd_amd.gen_er(2500, 2800, 9, seed=43, dtype=dtype))
    for A, B in (first, second):
        st, nnz, Cptr, Cidx, Cval = run_noalloc(A, B)
        assert st == 0, _lib.lib.dbd_last_error()
        C_noalloc = dbd_amd.CSR(Cptr.cpu().numpy(), Cidx[:nnz].cpu().numpy(), Cval[:nnz].cpu().numpy(), A.shape)
        assert C_noalloc.nnz == nnz
        C_batch = run_batch_entry(A, B)
        assert np.array_equal(C_noalloc.ptr, C_batch.ptr) and np.array_equal(C_noalloc.idx, C_batch.idx)
        ref = reference(A, B)
        compare(C_noalloc, ref, dtype, "no-alloc entry")
        compare(C_batch, ref, dtype, "batch entry")


def test_noalloc_refusals(gpu_ok):
    A = dbd_amd.gen_er(2000, 2000, 8, seed=44)
    st, nnz, _, _, _ = run_noalloc(A, None, capacity=10)
    assert st == 4 and nnz > 10  # DBD_ERR_BUDGET, with the size needed
    st2, nnz2, Cptr, Cidx, Cval = run_noalloc(A, None, capacity=nnz)  # exactly enough
    assert st2 == 0 and nnz2 == nnz
    compare(dbd_amd.CSR(Cptr.cpu().numpy(), Cidx.cpu().numpy(), Cval.cpu().numpy(), A.shape), reference(A),
            np.float64, "exact capacity")
    # workspace above the fixed array: refused before any device work
    big = int(_lib.lib.dbd_noalloc_scratch_bytes(1 << 24, 1 << 24))
    assert big > (256 << 20)
    v = dbd_amd.DeviceCSR.from_host(A).view()
    v.nrows = v.ncols = 1 << 24  # only the sizes are read before the refusal (B = A, so square)
    v.nnz = 1 << 24
    n = C.c_int64(-1)
    st3 = _lib.lib.dbd_spgemm_noalloc(C.byref(v), None, None, Cptr.data_ptr(), Cidx.data_ptr(), Cval.data_ptr(),
                                      nnz, C.byref(n))
    assert st3 == 4 and b"workspace" in _lib.lib.dbd_last_error()
