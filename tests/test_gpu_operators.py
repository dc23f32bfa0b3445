"""The C++ operator layer (vexq_amd/host) driven the way the reference's Go
tests drive exec.Operator: tests/cpp/test_exec.cpp, built by build(), run on
the GPU, one PASS/FAIL line per restated reference test."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "vexq_amd", "host", "test_exec")


def test_cpp_operator_suite():
    assert os.path.exists(EXE), "build the host layer first (make -C vexq_amd/host)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    fails = [l for l in r.stdout.splitlines() if l.startswith("FAIL")]
    assert r.returncode == 0 and not fails, "\n".join(fails) + r.stderr
    assert sum(l.startswith("PASS") for l in r.stdout.splitlines()) >= 11
