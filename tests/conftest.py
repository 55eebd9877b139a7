import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")
    # the autotuner's persistent cache stays out of the tests unless a test
    # names a file of its own (every autotune call then runs its trials)
    os.environ["FDSP_TUNE_CACHE"] = "off"
    # a fresh checkout has no built libraries: compile them once (the HIP
    # library cross-compiles for gfx950 without a GPU)
    if not os.path.exists(os.path.join(ROOT, "libfalcon_dsp_amd", "libfdsp_hip.so")):
        from libfalcon_dsp_amd import build
        build.build_all()
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)


_gpu_state = {}


def _gpu_available():
    if "ok" not in _gpu_state:
        # the HIP library must load; a missing library is an error, not a skip
        import ctypes

        import libfalcon_dsp_amd._native as N
        n = ctypes.c_int(0)
        rc = N.lib().fdsp_device_count(ctypes.byref(n))
        _gpu_state["ok"] = rc == 0 and n.value > 0
    return _gpu_state["ok"]


def pytest_runtest_setup(item):
    if item.get_closest_marker("gpu") is not None and not _gpu_available():
        pytest.skip("no HIP device visible")
