"""Runs the C++ API test binary (shell-encryption_amd/cpp/tests/rns_api_test),
which drives the rlwe:: classes on the GPU and checks them against the oracle."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "shell-encryption_amd", "cpp", "tests", "rns_api_test")

pytestmark = pytest.mark.gpu


def test_cpp_api_suite():
    assert os.path.exists(BIN), "rns_api_test is not built; run __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "checks passed" in r.stdout
