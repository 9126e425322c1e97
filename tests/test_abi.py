"""The C ABI: libdbd.so loads without a GPU, exports every function include/dbd.h declares,
and the Python declarations match the header. No device compute here."""
import os
import re
import subprocess

import pytest

import dbd_amd
from dbd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dbd.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(dbd_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_declares_api():
    fns = header_functions()
    assert len(fns) == len(_lib.SIGNATURES)
    for must in ("dbd_create", "dbd_run", "dbd_symbolic", "dbd_numeric", "dbd_backward", "dbd_batch_launch",
                 "dbd_autotune", "dbd_read_matrix", "dbd_plan_blocks"):
        assert must in fns


def test_every_declared_symbol_is_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.lib_path], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = [f for f in header_functions() if f not in exported]
    assert not missing, f"declared but not exported: {missing}"


def test_python_signatures_cover_header():
    assert sorted(_lib.SIGNATURES) == header_functions()


def test_abi_version_and_errors_without_gpu():
    assert _lib.lib.dbd_abi_version() == 2
    import torch
    if torch.cuda.is_available():
        pytest.skip("this check is for hosts without a GPU")
    with pytest.raises(dbd_amd.DBDError) as e:
        dbd_amd.Context()
    assert e.value.status == 2 and "no HIP device" in str(e.value)


def test_invalid_options_rejected():
    import ctypes as C
    o = _lib.dbd_opts()
    _lib.lib.dbd_default_opts(C.byref(o))
    o.block_rows = 1000  # not a multiple of 256
    h = C.c_void_p()
    st = _lib.lib.dbd_create(C.byref(o), C.byref(h))
    assert st == 1 and b"multiple of 256" in _lib.lib.dbd_last_error()


def test_cli_usage_on_bad_argv():
    exe = os.path.join(ROOT, "dbd_amd", "bin", "SSpMM")
    r = subprocess.run([exe, "a", "b", "c", "d", "e"], capture_output=True, text=True)
    assert r.returncode == 255
    assert "Invalid Arguments." in r.stdout and "Usage:" in r.stdout
    for name in ("compute", "d2h", "debug"):
        assert os.access(os.path.join(ROOT, "dbd_amd", "bin", name), os.X_OK)
