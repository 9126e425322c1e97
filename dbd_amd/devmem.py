"""Device-memory plumbing through torch (allocation and copies only; no compute)."""
from __future__ import annotations

import numpy as np


def copy_device_to_host(addr: int, nbytes: int, dtype) -> np.ndarray:
    """Copy `nbytes` at a device address into a new numpy array (via a torch staging tensor)."""
    import torch
    out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
    if nbytes == 0:
        return out
    src = torch.empty(0, dtype=torch.uint8, device="cuda")
    # wrap the raw pointer with a storage view: use cudart memcpy through torch.cuda
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    torch.cuda.synchronize()
    rc = hip.hipMemcpy(out.ctypes.data, ctypes.c_void_p(addr), nbytes, 2)  # hipMemcpyDeviceToHost
    if rc != 0:
        raise RuntimeError(f"hipMemcpy D2H failed ({rc})")
    del src
    return out
