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
    assert _lib.lib.dbd_abi_version() == 3
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


def test_candidate_tables_cover_every_knob_combination():
    """The autotuner's search space: 48 tables, every combination of the five knobs once, table 0
    the default (16 / 64 lanes, load 0.5, 8192-slot largest table, hub kernels at every width)."""
    n = dbd_amd.Context.table_count()
    assert n == 48
    seen = set()
    for t in range(n):
        k = dbd_amd.Context.table_knobs(t)
        seen.add((k["g_small"], k["g_big"], k["load"], k["top"], k["hub_min_cols"]))
    assert len(seen) == 48
    assert {x[0] for x in seen} == {16, 64} and {x[1] for x in seen} == {16, 64}
    assert {x[2] for x in seen} == {0.5, 0.25} and {x[3] for x in seen} == {4096, 8192}
    assert {x[4] for x in seen} == {0, 1 << 21, 2 ** 63 - 1}
    assert dbd_amd.Context.table_knobs(0) == {"g_small": 16, "g_big": 64, "load": 0.5, "top": 8192,
                                             "hub_min_cols": 0}
    with pytest.raises(dbd_amd.DBDError):
        dbd_amd.Context.table_knobs(48)
