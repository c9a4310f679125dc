"""The C++ facade (rampack_amd/host/rampack_gpu.hpp) driven like the reference's Simulation API."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "facade_test")


@pytest.mark.gpu
def test_cpp_facade_program():
    assert os.path.exists(BIN), "build() compiles tests/cpp/facade_test"
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=200)
    print(res.stdout, res.stderr)
    assert res.returncode == 0
    assert "facade ok" in res.stdout
