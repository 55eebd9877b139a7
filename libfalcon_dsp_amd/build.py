"""In-tree build of libfdsp_hip.so (gfx950) and the C++ drop-in library.

    python -m libfalcon_dsp_amd.build [--incremental]

Each HIP source compiles to its own object under libfalcon_dsp_amd/_build/
(in parallel), then the objects link into libfdsp_hip.so. --incremental
recompiles only objects older than their source or a csrc/ header.

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
               for f in ("fdsp_kernels.hip", "fdsp_chan_groups.hip", "fdsp_chan_seg.hip", "fdsp_capi.hip", "fdsp_serve.hip", "fdsp_fm.hip",
                         "fdsp_iir.hip", "fdsp_predefined.hip")]
PREDEFINED_INC = os.path.join(CSRC, "fdsp_predefined_table.inc")
HIP_LIB = os.path.join(HERE, "libfdsp_hip.so")
DROPIN_SOURCES_DIR = os.path.join(HERE, "src")
DROPIN_LIB = os.path.join(HERE, "libfalcon_dsp.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
OBJ_DIR = os.path.join(HERE, "_build")


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


def build_hip(incremental: bool = False):
    from concurrent.futures import ThreadPoolExecutor
    build_predefined_table()
    os.makedirs(OBJ_DIR, exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".h", ".inc"))]
    headers.append(os.path.join(ROOT, "include", "fdsp.h"))
    newest_header = max(os.path.getmtime(h) for h in headers)
    objs, todo = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ_DIR, os.path.basename(src).replace(".hip", ".o"))
        objs.append(obj)
        if (not incremental or not os.path.exists(obj)
                or os.path.getmtime(obj) < max(os.path.getmtime(src), newest_header)):
            todo.append((src, obj))
    jobs = max(1, min(len(todo), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16))
    with ThreadPoolExecutor(jobs) as pool:
        for f in [pool.submit(_run, [HIPCC, *HIP_FLAGS, "-c", src, "-o", obj]) for src, obj in todo]:
            f.result()
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", HIP_LIB])


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


def build_all(incremental: bool = False):
    """The product: the HIP library, the C++ drop-in library and the C++ test
    programs that drive it. The CPU checker under oracle/ is built by its own
    Makefile (make -C oracle), never by this module."""
    build_hip(incremental)
    build_dropin()
    build_dropin_test()


if __name__ == "__main__":
    build_all(incremental="--incremental" in sys.argv[1:])
    sys.exit(0)
