"""Sanitizer runs of the host code (CPU only; no GPU is used).

- ThreadSanitizer over the serving scheduler (fdsp_serve_sched.h), the
  queueing, batching and completion logic the server's worker and client
  threads share: tests/sched/test_sched.cc, built with g++ -fsanitize=thread.
- AddressSanitizer and UndefinedBehaviorSanitizer over the C ABI's host layer
  and the drop-in library: the library's sources compiled with the
  sanitizers on their host code, linked into tests/sanitize/test_capi_host.cc,
  which drives every path that needs no device (validation, errors, pure host
  functions, file I/O, constructor clean-up without a device).

A sanitizer report or a failed check ends the program with a non-zero status.
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "libfalcon_dsp_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
CLANGXX = "/opt/rocm/llvm/bin/clang++"


def _run(cmd, env=None, timeout=600):
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=timeout)
    assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stdout[-4000:]}\n{r.stderr[-4000:]}"
    return r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_scheduler_under_thread_sanitizer(tmp_path):
    exe = str(tmp_path / "test_sched")
    _run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "sched", "test_sched.cc"), "-o", exe, "-pthread"])
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1:second_deadlock_stack=1")
    out = _run([exe], env=env, timeout=300)
    assert "scheduler checks passed" in out


@pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANGXX)), reason="ROCm compilers not installed")
def test_host_layer_under_address_and_ub_sanitizers(tmp_path):
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    # host code sanitized, device code as built for the product
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]

    def compile_hip(f):
        obj = str(tmp_path / (f + ".o"))
        _run([HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC", "-Xclang", "-target-feature",
              "-Xclang", "-packed-fp32-ops"] + san + ["-c", os.path.join(CSRC, f), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_hip, srcs))
    inc = ["-I" + os.path.join(ROOT, "libfalcon_dsp_amd", "hdr"), "-I" + os.path.join(ROOT, "include")]
    host_san = ["-fsanitize=address,undefined"]
    dropin = str(tmp_path / "dropin.o")
    _run([CLANGXX, "-std=c++17", "-O1", "-g", "-fPIC"] + host_san + inc +
         ["-c", os.path.join(ROOT, "libfalcon_dsp_amd", "src", "falcon_dsp_dropin.cc"), "-o", dropin])
    drv = str(tmp_path / "drv.o")
    _run([CLANGXX, "-std=c++17", "-O1", "-g"] + host_san + inc +
         ["-c", os.path.join(ROOT, "tests", "sanitize", "test_capi_host.cc"), "-o", drv])
    exe = str(tmp_path / "capi_host")
    _run([CLANGXX] + host_san + [drv, dropin] + objs +
         ["-o", exe, "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    # the HIP runtime's own allocations outlive main: leak checking stays off
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    out = _run([exe, str(tmp_path)], env=env, timeout=300)
    assert "host-layer checks passed" in out
