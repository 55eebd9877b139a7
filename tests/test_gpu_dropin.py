"""The C++ drop-in classes (reference class names over the C ABI) on the GPU:
tests/dropin/test_dropin.cc, built by libfalcon_dsp_amd.build, drives them
like the reference's unit tests drive its classes (KATs, streaming, closed
forms, batched class against the per-channel class)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin", "test_dropin")


def test_dropin_program():
    assert os.path.exists(BIN), "build first: python -m libfalcon_dsp_amd.build"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
