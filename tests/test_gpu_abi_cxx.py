"""A C++ program that includes include/mpdg_heat.h and calls create / vmult / vmult_noalloc /
destroy through the C ABI (tests/cxx/abi_smoke.cpp, built by __graft_entry__.build())."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cxx_caller_of_the_c_abi(cuda):
    exe = os.path.join(ROOT, "tests", "cxx", "abi_smoke")
    assert os.path.exists(exe), "tests/cxx/abi_smoke is not built (run __graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "abi_smoke ok" in r.stdout
