"""Build libdbd.so (gfx950 HIP kernels + C++ engine) and the CLI drop-ins, in-tree.

Outputs:
    dbd_amd/libdbd.so                 the C-ABI library (include/dbd.h)
    dbd_amd/bin/{compute,d2h,SSpMM,debug}   reference-compatible command lines

Run: python dbd_amd/build.py   (the package itself loads libdbd.so, so the build runs as a script)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
BIN = os.path.join(HERE, "bin")
LIB = os.path.join(HERE, "libdbd.so")

ARCH = os.environ.get("DBD_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    # packed FP32 ops disabled on purpose (hipcc flag only)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-pass-failed", "-I", os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", "-I", os.path.join(ROOT, "include")]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{msg}")


def build(verbose: bool = True, lib_out: str | None = None, defines: list[str] | None = None,
          cli: bool = True) -> str:
    """Compile everything; `lib_out` / `defines` build an alternative library (e.g. a kernel
    tuning variant) without touching the default one."""
    lib = lib_out or LIB
    bdir = BUILD if lib_out is None else lib_out + ".build"
    os.makedirs(bdir, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    dflags = [f"-D{d}" for d in (defines or [])]
    objs = {
        "kernels.o": [HIPCC, *HIP_FLAGS, *dflags, "-c", os.path.join(CSRC, "kernels.hip")],
        "engine.o": [HIPCC, *HIP_FLAGS, *dflags, "-c", os.path.join(CSRC, "engine.hip")],
        "noalloc.o": [HIPCC, *HIP_FLAGS, *dflags, "-c", os.path.join(CSRC, "noalloc.hip")],
        "host.o": [CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, "host.cpp")],
    }
    jobs = []
    with ThreadPoolExecutor(max_workers=4) as ex:
        for name, cmd in objs.items():
            jobs.append(ex.submit(_run, cmd + ["-o", os.path.join(bdir, name)]))
        for j in jobs:
            j.result()
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp,
          *[os.path.join(bdir, n) for n in objs], "-lpthread"])
    os.replace(tmp, lib)
    if not cli:
        return lib
    # rocSPARSE comparison harness (a program of its own: bench and tests only)
    _run([HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++17", os.path.join(CSRC, "vendor_spgemm.cpp"), "-o",
          os.path.join(BIN, "rocsparse_spgemm"), "-L", HERE, "-ldbd", "-Wl,-rpath,$ORIGIN/..", "-L", "/opt/rocm/lib",
          "-lrocsparse", "-Wl,-rpath,/opt/rocm/lib"])
    modes = {"compute": ["-DDBD_CLI_MODE=0"], "d2h": ["-DDBD_CLI_MODE=1"],
             "SSpMM": ["-DDBD_CLI_MODE=2"], "debug": ["-DDBD_CLI_MODE=2", "-DDBD_CLI_DEBUG"]}
    with ThreadPoolExecutor(max_workers=4) as ex:
        futs = [ex.submit(_run, [CXX, *CXX_FLAGS, *defs, os.path.join(CSRC, "cli.cpp"), "-o",
                                 os.path.join(BIN, name), "-L", HERE, "-ldbd",
                                 "-Wl,-rpath,$ORIGIN/..", "-lpthread"])
                for name, defs in modes.items()]
        for f in futs:
            f.result()
    if verbose:
        print(f"built {LIB} and {BIN}/{{compute,d2h,SSpMM,debug,rocsparse_spgemm}} for {ARCH}")
    return LIB


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-out", default=None, help="write the library here instead of dbd_amd/libdbd.so")
    ap.add_argument("-D", dest="defines", action="append", default=[], help="extra preprocessor define")
    a = ap.parse_args()
    try:
        build(lib_out=a.lib_out, defines=a.defines, cli=a.lib_out is None)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
