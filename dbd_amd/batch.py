"""Batched form: many independent C_i = A_i * B_i as one block-diagonal product.

A_cat stacks the rows of every A_i with A_i's column indices shifted to B_i's first row in
B_cat; B_cat stacks the rows of every B_i with column indices shifted by the columns of the
members before it. Then C_cat = A_cat * B_cat holds every C_i in its own row range and column
range, and one plan (row analysis, categories, symbolic, numeric) covers all members: the
library captures that whole pipeline in one hipGraph and replays it with one launch.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .api import CSR, Context, DeviceCSR


@dataclass
class BatchLayout:
    row_off: np.ndarray   # first row of member i in A_cat / C_cat (len n+1)
    brow_off: np.ndarray  # first row of member i in B_cat
    col_off: np.ndarray   # first column of member i in B_cat / C_cat


def concat(As: list[CSR], Bs: list[CSR]) -> tuple[CSR, CSR, BatchLayout]:
    assert len(As) == len(Bs) and As
    dt = As[0].val.dtype
    row_off = np.zeros(len(As) + 1, np.int64)
    brow_off = np.zeros(len(As) + 1, np.int64)
    col_off = np.zeros(len(As) + 1, np.int64)
    for i, (A, B) in enumerate(zip(As, Bs)):
        assert A.shape[1] == B.shape[0], f"member {i}: inner dimensions differ"
        row_off[i + 1] = row_off[i] + A.shape[0]
        brow_off[i + 1] = brow_off[i] + B.shape[0]
        col_off[i + 1] = col_off[i] + B.shape[1]
    if brow_off[-1] >= 2 ** 31 or col_off[-1] >= 2 ** 31 or row_off[-1] >= 2 ** 31:
        raise OverflowError("batched dimensions must stay below 2^31")

    def stack(Ms, shift):
        nnz = np.cumsum([0] + [M.nnz for M in Ms])
        ptr = np.empty(sum(M.shape[0] for M in Ms) + 1, np.int64)
        idx = np.empty(int(nnz[-1]), np.int32)
        val = np.empty(int(nnz[-1]), dt)
        r = 0
        for i, M in enumerate(Ms):
            ptr[r:r + M.shape[0]] = M.ptr[:-1] + nnz[i]
            idx[nnz[i]:nnz[i + 1]] = M.idx + np.int32(shift[i])
            val[nnz[i]:nnz[i + 1]] = M.val
            r += M.shape[0]
        ptr[-1] = nnz[-1]
        return ptr, idx, val

    ap, ai, av = stack(As, brow_off)
    bp, bi, bv = stack(Bs, col_off)
    A_cat = CSR(ap, ai, av, (int(row_off[-1]), int(brow_off[-1])))
    B_cat = CSR(bp, bi, bv, (int(brow_off[-1]), int(col_off[-1])))
    return A_cat, B_cat, BatchLayout(row_off, brow_off, col_off)


def split(Cptr: np.ndarray, Cidx: np.ndarray, Cval: np.ndarray, lay: BatchLayout) -> list[CSR]:
    out = []
    for i in range(len(lay.row_off) - 1):
        r0, r1 = lay.row_off[i], lay.row_off[i + 1]
        e0, e1 = Cptr[r0], Cptr[r1]
        out.append(CSR(Cptr[r0:r1 + 1] - e0, Cidx[e0:e1] - np.int32(lay.col_off[i]), Cval[e0:e1],
                       (int(r1 - r0), int(lay.col_off[i + 1] - lay.col_off[i]))))
    return out


class BatchRunner:
    """Holds a context prepared for one batch; `launch()` replays the captured graph. C is
    allocated at the exact nnz that the preparation's symbolic pass reports. With tune=True the
    category table is autotuned on the concatenation first (`table` holds the chosen id, for
    running the plain per-matrix baseline with the same table)."""

    def __init__(self, A_cat: CSR, B_cat: CSR, dtype=None, block_rows: int = 4096, tune: bool = False,
                 table: int | None = None):
        import torch
        self.dA, self.dB = DeviceCSR.from_host(A_cat), DeviceCSR.from_host(B_cat)
        self.ctx = Context(block_rows=block_rows, dtype=dtype or A_cat.val.dtype)
        self.table = table
        if table is not None:
            self.ctx.set_table(table)
        elif tune:
            self.ctx.set_operands(self.dA, self.dB)
            self.table = self.ctx.autotune()
        self.nnz = self.ctx.batch_prepare(self.dA, self.dB)
        tdt = torch.float64 if A_cat.val.dtype == np.float64 else torch.float32
        self.Cptr = torch.empty(A_cat.shape[0] + 1, dtype=torch.int64, device="cuda")
        self.Cidx = torch.empty(max(self.nnz, 1), dtype=torch.int32, device="cuda")
        self.Cval = torch.empty(max(self.nnz, 1), dtype=tdt, device="cuda")

    def launch(self):
        self.ctx.batch_launch(self.Cptr, self.Cidx, self.Cval)

    def result(self):
        p = self.Cptr.cpu().numpy()
        n = int(p[-1])
        return p, self.Cidx[:n].cpu().numpy(), self.Cval[:n].cpu().numpy()

    def close(self):
        self.ctx.close()


def plain_each(As: list[CSR], Bs: list[CSR], block_rows: int = 4096, keep: bool = False,
               table: int | None = None):
    """Baseline: the plain pipeline once per member (plan, symbolic, numeric per matrix), with
    operands already on the device; `table` pins the category table (e.g. the batch's tuned
    one). Returns the list of device results if keep."""
    import torch
    dev = [(DeviceCSR.from_host(A), DeviceCSR.from_host(B)) for A, B in zip(As, Bs)]
    ctx = Context(block_rows=block_rows, dtype=As[0].val.dtype)
    if table is not None:
        ctx.set_table(table)
    tdt = torch.float64 if As[0].val.dtype == np.float64 else torch.float32

    def run():
        outs = []
        for (dA, dB), A in zip(dev, As):
            ctx.set_operands(dA, dB)
            ctx.plan()
            Cptr = torch.empty(A.shape[0] + 1, dtype=torch.int64, device="cuda")
            nnz = ctx.symbolic(Cptr)
            Cidx = torch.empty(max(nnz, 1), dtype=torch.int32, device="cuda")
            Cval = torch.empty(max(nnz, 1), dtype=tdt, device="cuda")
            ctx.numeric(Cptr, Cidx, Cval)
            outs.append((Cptr, Cidx, Cval) if keep else None)
        return outs

    return ctx, run
