"""Backward (masked gradients) on the GPU against the FP64 oracle gradient and central finite
differences, directly through the C ABI and through torch.autograd.

Tolerances: |g - g_ref| <= rtol * (the same gradient computed with |A|, |B|, |dC|), with
rtol = 1e-12 for FP64 and 1e-5 for FP32 (FP32 dB uses atomics, so summation order varies)."""
import numpy as np
import pytest
import scipy.sparse as sp

import dbd_amd
import oracle

pytestmark = pytest.mark.gpu
RT = {np.float64: 1e-12, np.float32: 1e-5}


def dev(M):
    return dbd_amd.DeviceCSR.from_host(M)


def run_backward(A, B, W_seed=0, block_rows=512):
    import torch
    dt = torch.float64 if A.val.dtype == np.float64 else torch.float32
    Ad, Bd = dev(A), (dev(B) if B is not A else None)
    with dbd_amd.Context(block_rows=block_rows, dtype=A.val.dtype) as ctx:
        ctx.set_operands(Ad, Bd if Bd is not None else Ad)
        ctx.plan()
        Cptr = torch.empty(A.shape[0] + 1, dtype=torch.int64, device="cuda")
        nnz = ctx.symbolic(Cptr)
        Cidx = torch.empty(nnz, dtype=torch.int32, device="cuda")
        Cval = torch.empty(nnz, dtype=dt, device="cuda")
        ctx.numeric(Cptr, Cidx, Cval)
        rng = np.random.default_rng(W_seed)
        W = torch.from_numpy(rng.standard_normal(nnz).astype(A.val.dtype)).cuda()
        gA = torch.empty(A.nnz, dtype=dt, device="cuda")
        gB = torch.zeros(B.nnz, dtype=dt, device="cuda")
        ctx.backward(Cptr, Cidx, W, gA, gB)
        torch.cuda.synchronize()
        return Cptr.cpu().numpy(), Cidx.cpu().numpy(), W.cpu().numpy(), gA.cpu().numpy(), gB.cpu().numpy()


def check_grads(A, B, Cptr, Cidx, W, gA, gB):
    dA, dB, miss = oracle.backward(A, B, Cptr, Cidx, W.astype(np.float64))
    assert miss == 0
    absA = dbd_amd.CSR(A.ptr, A.idx, np.abs(A.val).astype(np.float64), A.shape)
    absB = dbd_amd.CSR(B.ptr, B.idx, np.abs(B.val).astype(np.float64), B.shape)
    sA, sB, _ = oracle.backward(absA, absB, Cptr, Cidx, np.abs(W).astype(np.float64))
    rt = RT[A.val.dtype.type]
    ea = np.abs(gA - dA) > rt * sA + 1e-300
    eb = np.abs(gB - dB) > rt * sB + 1e-300
    assert not ea.any(), f"dA: {ea.sum()} entries off"
    assert not eb.any(), f"dB: {eb.sum()} entries off"


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_backward_random(gpu_ok, dtype):
    rng = np.random.default_rng(1)
    A = dbd_amd.CSR.from_scipy(sp.random(3000, 2000, density=0.004, random_state=1, format="csr",
                                         data_rvs=lambda s: rng.standard_normal(s)), dtype)
    B = dbd_amd.CSR.from_scipy(sp.random(2000, 2500, density=0.006, random_state=2, format="csr",
                                         data_rvs=lambda s: rng.standard_normal(s)), dtype)
    check_grads(A, B, *run_backward(A, B))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_backward_all_categories(gpu_ok, dtype):
    """R-MAT rows span tiny tables to the long-row window path."""
    A = dbd_amd.gen_rmat(13, 16, seed=21, add_identity=True, row_normalize=True, dtype=dtype)
    B = dbd_amd.gen_rmat(13, 16, seed=22, dtype=dtype)
    Cptr, Cidx, W, gA, gB = run_backward(A, B)
    check_grads(A, B, Cptr, Cidx, W, gA, gB)


def test_backward_long_rows(gpu_ok):
    rng = np.random.default_rng(3)
    n = 20000
    M = sp.random(n, n, density=0.0004, random_state=3, format="lil")
    for r, d in ((5, 4000), (77, 9000)):
        M[r, rng.choice(n, d, replace=False)] = rng.standard_normal(d)
    A = dbd_amd.CSR.from_scipy(M.tocsr())
    B = dbd_amd.CSR.from_scipy(sp.random(n, n, density=0.002, random_state=4, format="csr"))
    Cptr, Cidx, W, gA, gB = run_backward(A, B)
    assert np.diff(Cptr).max() > 4096  # the window path ran
    check_grads(A, B, Cptr, Cidx, W, gA, gB)


def test_autograd_function_fd_fp64(gpu_ok):
    """torch.autograd over SpGEMMFunction; loss = sum(W * C.values); central differences with
    h = 1e-3 on 64 random entries of A and of B (the loss is linear in each entry)."""
    import torch
    from dbd_amd.autograd import spgemm_values
    rng = np.random.default_rng(4)
    A = dbd_amd.CSR.from_scipy(sp.random(500, 400, density=0.02, random_state=5, format="csr",
                                         data_rvs=lambda s: rng.standard_normal(s)))
    B = dbd_amd.CSR.from_scipy(sp.random(400, 450, density=0.02, random_state=6, format="csr",
                                         data_rvs=lambda s: rng.standard_normal(s)))
    Ad, Bd = dev(A), dev(B)
    av = Ad.val.clone().requires_grad_(True)
    bv = Bd.val.clone().requires_grad_(True)
    Cv, Cp, Ci = spgemm_values(Ad.ptr, Ad.idx, av, A.shape, Bd.ptr, Bd.idx, bv, B.shape, block_rows=256)
    W = torch.randn(Cv.shape, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(0))
    (Cv * W).sum().backward()
    ga, gb = av.grad.cpu().numpy(), bv.grad.cpu().numpy()

    def loss(a, b):
        with torch.no_grad():
            c, p, i = spgemm_values(Ad.ptr, Ad.idx, a, A.shape, Bd.ptr, Bd.idx, b, B.shape, block_rows=256)
            assert torch.equal(i, Ci)
            return float((c * W).sum())

    h = 1e-3
    for p in rng.choice(A.nnz, 64, replace=False):
        e = torch.zeros_like(av)
        e[p] = h
        fd = (loss(av.detach() + e, bv.detach()) - loss(av.detach() - e, bv.detach())) / (2 * h)
        assert abs(fd - ga[p]) <= 1e-8 * max(1.0, abs(fd)), (p, fd, ga[p])
    for p in rng.choice(B.nnz, 64, replace=False):
        e = torch.zeros_like(bv)
        e[p] = h
        fd = (loss(av.detach(), bv.detach() + e) - loss(av.detach(), bv.detach() - e)) / (2 * h)
        assert abs(fd - gb[p]) <= 1e-8 * max(1.0, abs(fd)), (p, fd, gb[p])
    check_grads(A, B, Cp.cpu().numpy(), Ci.cpu().numpy(), W.cpu().numpy(), ga, gb)


def test_autograd_a_times_a(gpu_ok):
    """C = A*A with one tensor in both slots: autograd sums both gradient contributions."""
    import torch
    from dbd_amd.autograd import spgemm_values
    A = dbd_amd.gen_rmat(10, 8, seed=9)
    Ad = dev(A)
    av = Ad.val.clone().requires_grad_(True)
    Cv, Cp, Ci = spgemm_values(Ad.ptr, Ad.idx, av, A.shape, Ad.ptr, Ad.idx, av, A.shape, block_rows=256)
    Cv.sum().backward()
    dA, dB, _ = oracle.backward(A, A, Cp.cpu().numpy(), Ci.cpu().numpy(), np.ones(Ci.numel()))
    np.testing.assert_allclose(av.grad.cpu().numpy(), dA + dB, rtol=1e-12, atol=1e-12)


def test_autograd_rmat_fp32_vs_fp64_gradient(gpu_ok):
    """The differentiated configuration's shape at reduced scale (R-MAT scale 14 for both):
    A = row-normalised adjacency + self loops, B independent; FP32 gradients vs the FP64 oracle."""
    import torch
    from dbd_amd.autograd import spgemm_values
    A = dbd_amd.gen_rmat(14, 16, seed=1, add_identity=True, row_normalize=True, dtype=np.float32)
    B = dbd_amd.gen_rmat(14, 16, seed=2, dtype=np.float32)
    Ad, Bd = dev(A), dev(B)
    av = Ad.val.clone().requires_grad_(True)
    bv = Bd.val.clone().requires_grad_(True)
    Cv, Cp, Ci = spgemm_values(Ad.ptr, Ad.idx, av, A.shape, Bd.ptr, Bd.idx, bv, B.shape, block_rows=512)
    W = torch.randn(Cv.shape, dtype=torch.float32, device="cuda")
    (Cv * W).sum().backward()
    check_grads(A, B, Cp.cpu().numpy(), Ci.cpu().numpy(), W.cpu().numpy(), av.grad.cpu().numpy(),
                bv.grad.cpu().numpy())
