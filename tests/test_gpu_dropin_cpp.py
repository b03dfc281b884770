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


CHECK_BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "check_dropin.bin")


def test_cpp_check_spo_wfc_distance_tables():
    """check_spo / check_wfc / distance tables (tests/cpp/check_dropin.cpp): SPOSetHIP, DistanceTableHIP
    (AA, BA), WaveFunctionBatchHIP (flex_*) and WaveFunctionHIP against the FP64 oracle, in the manner
    of the reference's src/Drivers/check_spo.cpp:205-347 and check_wfc.cpp:190-485."""
    assert os.path.exists(CHECK_BIN), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    out = subprocess.run([CHECK_BIN], capture_output=True, text=True, timeout=300)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "All checks passed" in out.stdout
