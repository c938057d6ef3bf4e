"""The C++ drop-in headers (include/ubg/*.hpp): code written against the
reference's Simulation / MG / Single2DGrid API compiles against them with g++
(CPU) and runs on the GPU matching the oracle (GPU)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "_build", "dropin_test")


def test_dropin_headers_compile_and_link():
    out = subprocess.run(["make", "-B", "-C", CPP], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert os.path.exists(BIN)


@pytest.mark.gpu
def test_dropin_program_runs_on_gpu(gpu):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/_build/dropin_test missing: run __graft_entry__.build()")
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ALL OK" in out.stdout
