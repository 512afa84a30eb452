"""The C++ host adapter (include/dolfin_amd/HipAssembler.hpp) driving the library from a
compiled program, checked against the reference's known-answer norms."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "_build", "cpp_system_assembler")


def test_cpp_adapter_known_answers():
    assert os.path.exists(BIN), "build the library first (python -m dolfin_amd.build)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "adapter checks passed" in r.stdout
