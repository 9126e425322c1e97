"""dbd_amd - MI355X-native Dynamic Block Distributor (CSR SpGEMM, C = A*B).

Thin Python binding over the C ABI in include/dbd.h (libdbd.so, built in-tree by
``python dbd_amd/build.py``). There is no Python or CPU fallback of the product: if the
library is missing, importing the binding raises.

Main entry points:
    Context            the reference pipeline (plan / run in compute, d2h or io mode)
    spgemm(A, B, ...)  one-shot convenience over Context
    CSR                host CSR container (numpy arrays, int64 row pointers)
    gen_er / gen_web / gen_rmat / gen_skewed   seeded synthetic generators (host, C++)
    read_matrix / write_outputs / plan_blocks / analyze_block   host-side file + plan logic
    SpGEMMFunction     torch.autograd.Function (dbd_amd.autograd)
"""
from __future__ import annotations

from ._lib import (DBDError, MODE_COMPUTE, MODE_D2H, MODE_IO, F32, F64, lib, lib_path)  # noqa: F401
from .api import (CSR, Context, DeviceCSR, analyze_block, gen_er, gen_rmat, gen_skewed, gen_web,  # noqa: F401
                  plan_blocks, read_matrix, spgemm, write_csr_cache, write_outputs)

__all__ = [
    "DBDError", "MODE_COMPUTE", "MODE_D2H", "MODE_IO", "F32", "F64", "lib", "lib_path",
    "CSR", "Context", "DeviceCSR", "analyze_block", "gen_er", "gen_rmat", "gen_skewed", "gen_web",
    "plan_blocks", "read_matrix", "spgemm", "write_csr_cache", "write_outputs",
]
