"""The C++ API (include/unap_hip.hpp) driven like the reference's examples."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "unap_amd", "lib", "api_example")


def test_cpp_api_example():
    assert os.path.exists(EXE), "build with __graft_entry__.build()"
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("OK")
