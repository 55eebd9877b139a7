"""In-tree build of libfdsp_hip.so (gfx950) and the C++ drop-in library.

    python -m libfalcon_dsp_amd.build

Packed FP32 instructions are switched off for device code
(-target-feature -packed-fp32-ops): the kernels are written for scalar FP32
FMAs only.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIP_SOURCES = [os.path.join(CSRC, f)
               for f in ("fdsp_kernels.hip", "fdsp_chan_groups.hip", "fdsp_capi.hip", "fdsp_serve.hip", "fdsp_fm.hip",
                         "fdsp_iir.hip", "fdsp_predefined.hip")]
PREDEFINED_INC = os.path.join(CSRC, "fdsp_predefined_table.inc")
HIP_LIB = os.path.join(HERE, "libfdsp_hip.so")
DROPIN_SOURCES_DIR = os.path.join(HERE, "src")
DROPIN_LIB = os.path.join(HERE, "libfalcon_dsp.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the device-only feature; that note is noise
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_predefined_table():
    """The predefined resampling filters, designed by libfalcon_dsp_amd/predefined.py."""
    from . import predefined
    n = predefined.write_include(PREDEFINED_INC)
    print(f"+ designed {n} predefined filters -> {os.path.relpath(PREDEFINED_INC, ROOT)}", flush=True)


def build_hip():
    build_predefined_table()
    _run([HIPCC, *HIP_FLAGS, *HIP_SOURCES, "-o", HIP_LIB])


def build_dropin():
    """C++ drop-in classes (reference class names and signatures) over the C ABI."""
    if not os.path.isdir(DROPIN_SOURCES_DIR):
        return
    srcs = sorted(os.path.join(dp, f) for dp, _, fs in os.walk(DROPIN_SOURCES_DIR) for f in fs if f.endswith(".cc"))
    if not srcs:
        return
    _run(["g++", "-O2", "-std=c++14", "-fPIC", "-shared", "-Wall", f"-I{os.path.join(HERE, 'hdr')}",
          f"-I{os.path.join(ROOT, 'include')}", *srcs, "-o", DROPIN_LIB, f"-L{HERE}", "-lfdsp_hip",
          "-Wl,-rpath,$ORIGIN"])


DROPIN_TEST_SRC = os.path.join(ROOT, "tests", "dropin", "test_dropin.cc")
DROPIN_TEST_BIN = os.path.join(ROOT, "tests", "dropin", "test_dropin")
FILES_TEST_SRC = os.path.join(ROOT, "tests", "dropin", "test_files.cc")
FILES_TEST_BIN = os.path.join(ROOT, "tests", "dropin", "test_files")


def build_dropin_test():
    """The C++ programs that drive the drop-in classes like the reference's unit
    tests (GPU), and its sample-file functions (CPU)."""
    if not os.path.exists(DROPIN_LIB):
        return
    for src, out in ((DROPIN_TEST_SRC, DROPIN_TEST_BIN), (FILES_TEST_SRC, FILES_TEST_BIN)):
        if os.path.exists(src):
            _run(["g++", "-O2", "-std=c++14", "-Wall", f"-I{os.path.join(HERE, 'hdr')}",
                  f"-I{os.path.join(ROOT, 'include')}", src, "-o", out, f"-L{HERE}", "-lfalcon_dsp", "-lfdsp_hip",
                  "-Wl,-rpath,$ORIGIN/../../libfalcon_dsp_amd"])


def build_oracle():
    """The CPU checker (oracle/) and, where the reference sources exist, oracle/_ref."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"])


def build_all():
    build_hip()
    build_dropin()
    build_dropin_test()
    build_oracle()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
