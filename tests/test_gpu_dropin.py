"""Runs the C++ program written against the drop-in hala:: headers (built by
__graft_entry__.build() with plain g++) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_dropin")


def test_cpp_dropin_program():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/test_dropin is not built: run __graft_entry__.build()")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
