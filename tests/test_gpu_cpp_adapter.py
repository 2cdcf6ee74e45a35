"""The C++ drop-in adapter (include/oceano_operator.hpp) run from a compiled program."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_adapter")


def test_cpp_adapter_time_loop(cuda):
    assert os.path.exists(BIN), "build the C++ test with oceano_amd/build.py"
    print("[cpp] running %s" % BIN, flush=True)
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert out.stdout.count("bitwise identical") == 3 and "OK" in out.stdout
