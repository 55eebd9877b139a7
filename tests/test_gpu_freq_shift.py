"""NCO frequency shift on the GPU vs the oracle (reference
src/transform/falcon_dsp_freq_shift.cc:164-192, _cuda.cu:355-488).

Bound: max |err| <= 1e-4 * RMS(x) against the oracle; the carried index
(samples handled mod R) must equal the oracle's exactly after every chunk.
The deviation from the exact FP64 closed form x*exp(j w ((s+n) mod R)) is
checked too (same bound), and the deviation from an ideal rollover-free NCO is
reported by the bench, not asserted (the reference restarts at R by design).
"""
import numpy as np
import pytest

import libfalcon_dsp_amd as fd
from libfalcon_dsp_amd.signals import cwgn
from libfalcon_dsp_amd.workloads import Nco16
from oracle import oracle as O
from tests.helpers import assert_close_rms, chunks, rms

pytestmark = pytest.mark.gpu


def test_single_shift_kat_sizes():
    fs = 1_000_000
    for f in (23_000, -23_000, 1, 499_999, 0):
        x = cwgn(21, 4099)
        got = fd.FreqShift(fs, f).apply(x)
        want = O.FreqShift(fs, f).apply(x) if f else x
        assert_close_rms(got, want, rms(x), what=f"f={f}")


def test_nco_16_shift_chunked_against_oracle():
    w = Nco16()
    n = 1 << 20  # a slice of the 2^24-sample workload, same chunk size
    x = cwgn(1, n)
    shifts = w.shifts()
    gpu = fd.FreqShift(w.fs, shifts)
    refs = [O.FreqShift(w.fs, f) for f in shifts]
    xr = rms(x)
    for lo, hi in chunks(n, [w.chunk, 77_777, 3]):
        outs = gpu.apply(x[lo:hi])
        for c, f in enumerate(shifts):
            want = refs[c].apply(x[lo:hi])
            assert_close_rms(outs[c], want, xr, what=f"shift {f} chunk {lo}")
            assert gpu.samples_handled(c) == int(refs[c].samples_handled)
    # the exact FP64 closed form (with rollover), whole stream
    gpu.reset_state()
    outs = gpu.apply(x)
    for c, f in enumerate(shifts):
        want = O.fp64_freq_shift(x, w.fs, f)
        assert_close_rms(outs[c], want, xr, what=f"fp64 shift {f}")


def test_rollover_inside_chunk_and_reset():
    fs, f = 1_000_000, 23_000  # R = 434783: the index restarts inside the stream
    x = cwgn(5, 1_000_003)
    gpu = fd.FreqShift(fs, f)
    y1 = gpu.apply(x)
    assert_close_rms(y1, O.FreqShift(fs, f).apply(x), rms(x))
    R, _ = O.freq_shift_params(fs, f)
    assert gpu.samples_handled() == x.size % R
    gpu.reset_state()
    assert gpu.samples_handled() == 0
    y2 = gpu.apply(x)
    np.testing.assert_array_equal(y1, y2)


def test_device_form_matches_host_form():
    fs = 25_000_000
    shifts = [-10_000_000, 3_333_333, 9_999_999]
    x = cwgn(8, 300_001)
    host = fd.FreqShift(fs, shifts).apply(x)
    dev = fd.FreqShift(fs, shifts)
    d_in = fd.DeviceBuffer.from_array(x)
    d_out = fd.DeviceBuffer(3 * x.size)
    dev.apply_device(d_in, x.size, d_out, x.size)
    fd.device_sync()
    got = d_out.download().reshape(3, x.size)
    for c in range(3):
        np.testing.assert_array_equal(got[c], host[c])


def test_empty_input_is_an_error():
    with pytest.raises(fd.FdspError):
        fd.FreqShift(1_000_000, 1000).apply(np.zeros(0, np.complex64))


def test_apply_chunks_equals_per_chunk_calls():
    # 16 chunks of 2^14 through 16 shifts in one launch
    # (fdsp_freq_shift_apply_chunks_device) equal 16 per-chunk calls
    w = Nco16(n=1 << 18, chunk=1 << 14)
    x = w.input()
    b = fd.FreqShift(w.fs, w.shifts())
    d_in = fd.DeviceBuffer.from_array(x)
    oa = fd.DeviceBuffer(16 * w.n)
    ob = fd.DeviceBuffer(16 * w.n)
    s = fd.Stream()
    a2 = fd.FreqShift(w.fs, w.shifts())
    for k in range(w.n // w.chunk):
        tmp = fd.DeviceBuffer(16 * w.chunk)
        a2.apply_device(d_in, w.chunk, tmp, w.chunk, s, in_offset=k * w.chunk)
        s.sync()
        t = tmp.download(16 * w.chunk).reshape(16, w.chunk)
        for c in range(16):
            oa.upload(t[c], offset=c * w.n + k * w.chunk)
    b.apply_chunks_device(d_in, w.chunk, w.n // w.chunk, ob, w.n, s)
    s.sync()
    np.testing.assert_array_equal(ob.download(16 * w.n), oa.download(16 * w.n))
    for c in range(16):
        assert b.samples_handled(c) == a2.samples_handled(c)


def test_set_state_continues_the_stream():
    # a handle set to another's position continues its index exactly
    fs = 1_000_000
    x = cwgn(44, 600_000)
    a, ref = fd.FreqShift(fs, 23_000), O.FreqShift(fs, 23_000)
    a.apply(x[:434_000])
    ref.apply(x[:434_000])
    b = fd.FreqShift(fs, 23_000)
    b.set_state(434_000)
    assert b.samples_handled() == a.samples_handled()
    ya, yb, r = a.apply(x[434_000:]), b.apply(x[434_000:]), ref.apply(x[434_000:])
    np.testing.assert_array_equal(yb, ya)
    assert_close_rms(yb, r, rms(x))
    assert b.samples_handled() == int(ref.samples_handled)
