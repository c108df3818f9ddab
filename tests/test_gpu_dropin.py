"""The C++ drop-in (include/mpstools_amd.hpp) on the GPU against the oracle:
runs tests/cpp/_build/test_dropin (built by __graft_entry__.build())."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "test_dropin")


def test_cpp_dropin_matches_oracle(gpu):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/_build/test_dropin not built (run __graft_entry__.build())")
    print("[case] C++ drop-in", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASS" in r.stdout
