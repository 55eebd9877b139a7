"""The C ABI library (libfalcon_dsp_amd/libfdsp_hip.so) on a machine without
a GPU: it loads, exports every function include/fdsp.h declares, and its
host-only entry points (rate helpers, version, error text) agree with the
oracle and the reference's known answers. Nothing here launches a kernel."""
import ctypes as C
import json
import os
import re

import numpy as np

import libfalcon_dsp_amd._native as N
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KAT = json.load(open(os.path.join(ROOT, "tests", "golden", "kat_reference.json")))


def _declared():
    text = open(os.path.join(ROOT, "include", "fdsp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fdsp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    assert os.path.exists(N.LIB_PATH), "build first: python -m libfalcon_dsp_amd.build"
    lib = C.CDLL(N.LIB_PATH)
    names = _declared()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    # and the Python binding declares a signature for each of them
    bound = set(N.SYMBOLS)
    assert not [n for n in names if n not in bound], "binding lacks declared functions"


def test_rate_helpers_through_the_abi():
    L = N.lib()
    for a, b, g in KAT["gcd"]["cases"]:
        assert L.fdsp_gcd(a, b) == g
    for a, b, m in KAT["lcm"]["cases"]:
        assert L.fdsp_lcm(a, b) == m
    n, d = C.c_int64(), C.c_int64()
    for f, md, nn, dd in KAT["rat_approx"]["cases"]:
        L.fdsp_rat_approx(f, md, C.byref(n), C.byref(d))
        assert (n.value, d.value) == (nn, dd)
    u = np.random.default_rng(1)
    r, w = C.c_uint32(), C.c_double()
    for _ in range(300):
        a, b = (int(v) for v in u.integers(1, 2**40, 2))
        assert L.fdsp_gcd(a, b) == O.gcd(a, b)
        f, md = float(u.uniform(0, 5)), int(u.integers(1, 5000))
        L.fdsp_rat_approx(f, md, C.byref(n), C.byref(d))
        assert (n.value, d.value) == O.rat_approx(f, md), (f, md)
        fs, fh = int(u.integers(1000, 2**31)), int(u.integers(-2**30, 2**30))
        assert L.fdsp_freq_shift_params(fs, fh, C.byref(r), C.byref(w)) == 0
        if fh:
            assert (r.value, w.value) == O.freq_shift_params(fs, fh)


def test_version_and_error_text():
    L = N.lib()
    assert L.fdsp_version() >= 100
    assert isinstance(L.fdsp_last_error(), bytes)
    r, w = C.c_uint32(), C.c_double()
    assert L.fdsp_freq_shift_params(0, 5, C.byref(r), C.byref(w)) == N.FDSP_EINVAL
    assert b"fs" in L.fdsp_last_error() or len(L.fdsp_last_error()) > 0


def test_serve_argument_checks_before_any_device_work():
    # the serving entry points validate their arguments before touching HIP
    L = N.lib()
    h = C.c_void_p()
    bad = [
        N.ServeConfig(0, 8, 1024, 4096, 0),           # fs 0
        N.ServeConfig(25_000_000, 0, 1024, 4096, 0),  # no streams
        N.ServeConfig(25_000_000, 8, 8192, 4096, 0),  # max_batch < max_chunk
    ]
    for cfg in bad:
        assert L.fdsp_serve_create(C.byref(cfg), C.byref(h)) == N.FDSP_EINVAL
        assert not h.value
    big = N.ServeConfig(25_000_000, 8, 1024, 1 << 31, 0)
    assert L.fdsp_serve_create(C.byref(big), C.byref(h)) == N.FDSP_ERANGE
    assert L.fdsp_serve_create(None, C.byref(h)) == N.FDSP_EINVAL
    res = N.ServeResult()
    assert L.fdsp_serve_poll(None, 1, 0, C.byref(res)) == N.FDSP_ENOTINIT
    assert L.fdsp_serve_release(None, 0) == N.FDSP_ENOTINIT
    assert L.fdsp_serve_destroy(None) == N.FDSP_OK
    # the ctypes mirrors of the header's structs
    assert C.sizeof(N.ServeConfig) == 32 and C.sizeof(N.ServeResult) == 64 and C.sizeof(N.ServeStats) == 72


def test_python_rate_helpers():
    import libfalcon_dsp_amd as fd
    assert fd.calculate_gcd(840, 210) == 210
    assert fd.calculate_lcm(382, 164) == 31324
    assert fd.rat_approx(0.768, 100) == (53, 69)
    # reduce_rates gives the channelizer's up/down from a rate pair
    assert fd.reduce_rates(20_000_000, 250_000) == (1, 80)
    assert fd.reduce_rates(1_000_000, 800_000) == (4, 5)


def test_dropin_library_exports_reference_classes():
    import subprocess
    lib = os.path.join(ROOT, "libfalcon_dsp_amd", "libfalcon_dsp.so")
    assert os.path.exists(lib), "build first: python -m libfalcon_dsp_amd.build"
    C.CDLL(lib)
    syms = subprocess.run(["nm", "-D", "-C", "--defined-only", lib], capture_output=True, text=True).stdout
    for name in ["falcon_dsp::falcon_dsp_freq_shift::apply", "falcon_dsp::falcon_dsp_freq_shift::get_freq_shift_params",
                 "falcon_dsp::falcon_dsp_freq_shift_cuda::initialize", "falcon_dsp::falcon_dsp_polyphase_resampler::apply",
                 "falcon_dsp::falcon_dsp_polyphase_resampler::needed_out_count",
                 "falcon_dsp::falcon_dsp_fir_filter::apply", "falcon_dsp::falcon_dsp_polar_discriminator::apply",
                 "falcon_dsp::falcon_dsp_multi_rate_channelizer::apply",
                 "falcon_dsp::falcon_dsp_multi_rate_channelizer_cuda::initialize", "falcon_dsp::resample(",
                 "falcon_dsp::upfirdn_cuda(", "falcon_dsp::fir_filter_cuda(", "falcon_dsp::polar_discriminator(",
                 "falcon_dsp::freq_shift_cuda(", "falcon_dsp::calculate_gcd", "falcon_dsp::rat_approx",
                 "falcon_dsp::falcon_dsp_fm_demodulator::demod_mono",
                 "falcon_dsp::falcon_dsp_fm_demodulator_cuda::initialize",
                 "falcon_dsp::falcon_dsp_fm_demodulator_cuda::demod_mono"]:
        assert name in syms, name
