"""CPU build of the library (librampack_host.so) for bounds checking: the kernel and host sources of
rampack_amd/csrc compiled by g++ against the host stand-in of the HIP runtime (include/hip/hip_runtime.h), with
AddressSanitizer.  Test infrastructure: nothing in the package loads it.

The one textual change to the sources: a kernel's dynamic LDS declaration (`extern __shared__ T name[];`) becomes a
pointer to the launch's LDS allocation, which has exactly the size the launch asks for.

    python tests/hostbuild/host_build.py [csrc_dir [library]]   # optional: another copy of rampack_amd/csrc, output path
"""
from __future__ import annotations

import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "rampack_amd", "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "librampack_host.so")

FLAGS = ["-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-ffp-contract=off", "-fsanitize=address",
         "-fno-omit-frame-pointer", "-w"]


def asan_runtime() -> str:
    return subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()


def host_kernel_source(src: str) -> str:
    text = open(src).read()
    text, n = re.subn(r"extern\s+__shared__\s+(\w+)\s+(\w+)\s*\[\s*\]\s*;",
                      r"\1 *\2 = static_cast<\1 *>(emu::dyn_lds());", text)
    if n == 0:
        raise RuntimeError(f"{src}: no dynamic LDS declaration found")
    out = os.path.join(OUT_DIR, "rp_kernels_host.cpp")
    with open(out, "w") as f:
        f.write(f'#line 1 "{src}"\n')
        f.write(text)
    return out


def build(csrc: str | None = None, lib: str = LIB) -> str:
    csrc = csrc or CSRC
    os.makedirs(OUT_DIR, exist_ok=True)
    ksrc = host_kernel_source(os.path.join(csrc, "rp_kernels.hip"))
    sources = [ksrc] + [os.path.join(csrc, s) for s in ("rp_capi.cpp", "rp_ramsnap.cpp", "rp_ramtrj.cpp")] + \
        [os.path.join(HERE, "lanes.cpp"), os.path.join(HERE, "mlp_stubs.cpp")]
    cmd = ["g++"] + FLAGS + ["-I", os.path.join(HERE, "include"), "-I", os.path.join(ROOT, "include"), "-I", csrc,
                             "-o", lib + ".tmp"] + sources
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("the CPU build failed")
    os.replace(lib + ".tmp", lib)
    return lib


if __name__ == "__main__":
    print(build(sys.argv[1] if len(sys.argv) > 1 else None, sys.argv[2] if len(sys.argv) > 2 else LIB))
