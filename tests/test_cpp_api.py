"""The C++ operator layer (include/ahead/column_operators/OperatorsAN.hpp) as a reference caller
uses it: tests/cpp/test_operators_an.cpp runs the Q1.1 and Q2.1 operator chains on device BATs and
compares every output with the oracle.  build() compiles the program."""
import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "tests" / "cpp" / "_build" / "test_operators_an"


def _need_exe():
    if not EXE.exists():
        pytest.skip("tests/cpp not built (run __graft_entry__.build())")


def test_cpp_program_links_without_gpu():
    """Every symbol the C++ headers use resolves against libahead_hip.so (no compute here)."""
    _need_exe()
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present: the gpu test runs the program")
    res = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 2 and "no GPU" in res.stdout, (res.returncode, res.stdout, res.stderr)


@pytest.mark.gpu
def test_cpp_operator_chains_match_oracle():
    _need_exe()
    env = dict(os.environ)
    res = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=300, env=env)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "PASSED" in res.stdout
