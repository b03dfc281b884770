"""Runs the C++ drop-in test (tests/cpp/test_dropin_dirac_det.cpp): the reference's DiracDeterminant
unit tests with qmcx::DelayedUpdateHIP as DU_TYPE.  The binary is built by __graft_entry__.build()."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_dropin_dirac_det.bin")


def test_cpp_dropin_dirac_det():
    assert os.path.exists(BIN), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "All tests passed" in out.stdout
