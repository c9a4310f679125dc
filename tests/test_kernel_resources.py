"""Compiler resource report for the device kernels (hipcc -Rpass-analysis=kernel-resource-usage, built here for
gfx950): no kernel may use private (scratch) memory.  A by-value kernel argument that the compiler fails to keep in
the kernel-argument segment is copied to scratch by every lane (about 1.2 KB per lane for the sweep parameters),
which multiplied the sweep's HBM writes a hundredfold when it happened."""
import os
import re
import shutil
import subprocess

import pytest

from rampack_amd import build

pytestmark = pytest.mark.skipif(shutil.which(build.hipcc()) is None and not os.path.exists(build.hipcc()),
                                reason="hipcc not available")


def test_no_kernel_uses_scratch(tmp_path):
    src = os.path.join(build.CSRC, "rp_kernels.hip")
    flags = [f for f in build.HIPCC_FLAGS if f not in ("-shared",)]
    cmd = [build.hipcc()] + flags + ["-I", os.path.join(build.ROOT, "include"), "-I", build.CSRC, "-c", src, "-o",
                                     str(tmp_path / "k.o"), "-Rpass-analysis=kernel-resource-usage"]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stderr[-2000:]
    names = re.findall(r"Function Name: (\S+)", res.stderr)
    scratch = [int(x) for x in re.findall(r"ScratchSize \[bytes/lane\]: (\d+)", res.stderr)]
    assert len(names) == len(scratch) and len(names) >= 12
    offenders = [(n, s) for n, s in zip(names, scratch) if s != 0]
    assert not offenders, offenders
