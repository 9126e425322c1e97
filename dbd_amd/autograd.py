"""torch.autograd.Function over the C ABI: C = A*B with fixed sparsity patterns and
differentiable values.

    Cval, Cptr, Cidx = spgemm_values(A_ptr, A_idx, A_val, A_shape, B_ptr, B_idx, B_val, B_shape)

Forward: dbd_plan + dbd_symbolic + dbd_numeric on torch's current stream.
Backward: dbd_backward, which reuses the forward's numeric row categories:
    dA = (dC * B^T) on pattern(A),   dB = (A^T * dC) on pattern(B).
No part of this runs on the CPU; tensors must be on the HIP device.
"""
from __future__ import annotations

import torch

from .api import Context, DeviceCSR


class SpGEMMFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a_val, b_val, a_ptr, a_idx, b_ptr, b_idx, a_shape, b_shape, block_rows):
        if not (a_val.is_cuda and b_val.is_cuda):
            raise RuntimeError("SpGEMMFunction needs device tensors (the HIP path has no CPU fallback)")
        if a_val.dtype != b_val.dtype or a_val.dtype not in (torch.float32, torch.float64):
            raise TypeError("A and B values must share dtype float32 or float64")
        a_val_d = a_val.detach().contiguous()
        b_val_d = b_val.detach().contiguous()
        A = DeviceCSR(a_ptr.contiguous(), a_idx.contiguous(), a_val_d, tuple(a_shape))
        same = (a_ptr.data_ptr() == b_ptr.data_ptr() and a_idx.data_ptr() == b_idx.data_ptr()
                and a_val.data_ptr() == b_val.data_ptr())
        B = A if same else DeviceCSR(b_ptr.contiguous(), b_idx.contiguous(), b_val_d, tuple(b_shape))
        dctx = Context(block_rows=block_rows, dtype=a_val.dtype == torch.float64 and "f64" or "f32",
                       stream=torch.cuda.current_stream())
        dctx.set_operands(A, B)
        dctx.plan()
        Cptr = torch.empty(int(a_shape[0]) + 1, dtype=torch.int64, device=a_val.device)
        nnz = dctx.symbolic(Cptr)
        Cidx = torch.empty(nnz, dtype=torch.int32, device=a_val.device)
        Cval = torch.empty(nnz, dtype=a_val.dtype, device=a_val.device)
        dctx.numeric(Cptr, Cidx, Cval)
        ctx.dctx = dctx
        ctx.keep = (A, B, a_val_d, b_val_d)
        ctx.save_for_backward(Cptr, Cidx)
        ctx.same = same
        ctx.mark_non_differentiable(Cptr, Cidx)
        return Cval, Cptr, Cidx

    @staticmethod
    def backward(ctx, g_val, g_ptr, g_idx):
        Cptr, Cidx = ctx.saved_tensors
        A, B, a_val, b_val = ctx.keep
        need_a, need_b = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if ctx.same:
            need_a = need_b = need_a or need_b
        dA = torch.empty_like(a_val) if need_a else None
        dB = torch.zeros_like(b_val) if need_b else None
        ctx.dctx.set_stream(torch.cuda.current_stream())
        ctx.dctx.backward(Cptr, Cidx, g_val.contiguous(), dA, dB)
        # C = A*A: both operand slots are the same tensor and autograd sums the two returns
        return dA, dB, None, None, None, None, None, None, None


def spgemm_values(a_ptr, a_idx, a_val, a_shape, b_ptr, b_idx, b_val, b_shape, block_rows: int = 4096):
    """Differentiable C = A*B. Returns (C values, C row pointer, C column indices)."""
    return SpGEMMFunction.apply(a_val, b_val, a_ptr, a_idx, b_ptr, b_idx, tuple(a_shape), tuple(b_shape),
                                block_rows)
