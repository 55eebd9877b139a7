"""Polyphase resampler on the GPU vs the oracle and the reference's KATs
(reference src/transform/falcon_dsp_polyphase_resampler.cc:114-346;
test/tests/transform/falcon_dsp_polyphase_resample_{cpp,cuda}_unit_tests:165-301).

Output counts and the carried (coeff_phase, xOffset) must match the oracle
exactly after every chunk; samples within 1e-4 * RMS(x).
"""
import numpy as np
import pytest

import libfalcon_dsp_amd as fd
from libfalcon_dsp_amd.signals import cwgn, kaiser_sinc
from libfalcon_dsp_amd.workloads import Resampler1ch
from oracle import oracle as O
from tests.helpers import assert_close_rms, chunks, rms

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize(
    "h,x,up,down,expected",
    [
        # upfirdn([.5,1,2], ones(7), 1, 2) = [.5, 3.5, 3.5, 3.5, 2]  (cpp _001, cuda _001)
        ([0.5, 1.0, 2.0], [1.0] * 7, 1, 2, [0.5, 3.5, 3.5, 3.5, 2.0]),
        # ones(1000), 1/2 -> 0.5 then 3.5 ...  (cuda _002)
        ([0.5, 1.0, 2.0], [1.0] * 1000, 1, 2, [0.5] + [3.5] * 500),
        # 7-tap filter, 1/2 (cuda _003)
        ([0.5, 0.5, 1.0, 0.3, 0.07, 0.1, 0.4], [1.0] * 1000, 1, 2, [0.5, 2.0, 2.37] + [2.87] * 498 + [0.87, 0.5]),
        # up 4 down 5 on ten ones (cuda _004)
        ([0.5, 1.0, 2.0], [1.0] * 10, 4, 5, [0.5, 1.0, 2.0, 0.0, 0.5, 1.0, 2.0, 0.0]),
    ],
)
def test_reference_kats(h, x, up, down, expected):
    r = fd.PolyphaseResampler(up, down, np.asarray(h, np.complex64))
    y = r.apply(np.asarray(x, np.complex64))
    # the reference emits ceil(N*up/down) outputs, a prefix of scipy's upfirdn
    assert y.size == -(-len(x) * up // down)
    n = min(y.size, len(expected))
    np.testing.assert_allclose(y[:n].real, expected[:n], atol=1e-5)
    np.testing.assert_allclose(y[:n].imag, 0.0, atol=1e-6)


def _stream_compare(up, down, taps, x, sizes):
    gpu = fd.PolyphaseResampler(up, down, taps)
    ref = O.Resampler(up, down, taps)
    xr = rms(x)
    for lo, hi in chunks(x.size, sizes):
        assert gpu.needed_out_count(hi - lo) == ref.needed_out_count(hi - lo)
        y = gpu.apply(x[lo:hi])
        r = ref.apply(x[lo:hi])
        assert y.size == r.size, f"chunk [{lo},{hi}): {y.size} vs {r.size}"
        assert_close_rms(y, r, xr, what=f"chunk [{lo},{hi})")
        ph, xo, it, ot = gpu.state()
        assert (ph, xo) == ref.state_fields, f"state after [{lo},{hi})"
        assert it == hi


def test_workload_resampler_1ch_chunked():
    w = Resampler1ch()
    x = w.input()
    taps = w.taps()
    gpu = fd.PolyphaseResampler(w.up, w.down, taps)
    ref = O.Resampler(w.up, w.down, taps)
    xr = rms(x)
    counts = []
    for lo, hi in chunks(x.size, [w.chunk]):
        y = gpu.apply(x[lo:hi])
        r = ref.apply(x[lo:hi])
        counts.append(y.size)
        assert y.size == r.size
        assert_close_rms(y, r, xr)
        assert gpu.state()[:2] == ref.state_fields
    assert sum(counts) == 838_861
    assert set(counts) == {52_428, 52_429}


@pytest.mark.parametrize("up,down,ntaps", [(1, 2, 7), (3, 2, 30), (2, 3, 31), (4, 5, 96), (1, 8, 32),
                                           (1, 80, 32), (1, 1600, 32), (5, 1, 17), (7, 1000, 70)])
def test_random_chunking_matches_oracle(up, down, ntaps):
    taps = kaiser_sinc(ntaps, 0.5 / max(up, down), 8.0, gain=up).astype(np.complex64)
    x = cwgn(up * 100 + down, 40_000)
    # chunks shorter than cpp-1, single samples, and chunks far shorter than down/up
    _stream_compare(up, down, taps, x, [1, 2, 3, 5, 997, 4096, 13, 1])


@pytest.mark.parametrize("up,down,ntaps", [(4, 5, 96), (1, 2, 32), (2, 3, 31), (3, 2, 30), (5, 1, 17), (5, 2, 80),
                                           (2, 9, 48)])
def test_long_chunks_match_oracle(up, down, ntaps):
    # long chunks in many tiles, ragged last tiles and the history carried in
    # between; 16, 24 and 32 taps per phase take the unrolled builds, and
    # (5, 2, 80) and (2, 9, 48) add an odd up and a decimation by 9 to them
    taps = kaiser_sinc(ntaps, 0.5 / max(up, down), 8.0, gain=up).astype(np.complex64)
    x = cwgn(17 * up + down, 1_100_001)
    _stream_compare(up, down, taps, x, [600_001, 500_000])


def test_chunk_invariance_and_reset():
    up, down = 3, 7
    taps = kaiser_sinc(45, 0.5 / 7, 8.0, gain=3).astype(np.complex64)
    x = cwgn(9, 30_000)
    r = fd.PolyphaseResampler(up, down, taps)
    whole = r.apply(x)
    r.reset_state()
    parts = np.concatenate([r.apply(x[lo:hi]) for lo, hi in chunks(x.size, [11, 1234, 7])])
    assert whole.size == parts.size
    assert_close_rms(parts, whole, rms(x), tol=1e-6)
    # against the FP64 brute force (zero-stuff, convolve, decimate)
    want = O.fp64_resample(x, taps, up, down)
    assert_close_rms(whole, want, rms(x))


def test_device_form_and_empty_chunk():
    up, down = 1, 20
    taps = kaiser_sinc(32, 0.5 / 20, 8.0).astype(np.complex64)
    x = cwgn(4, 100_000)
    r = fd.PolyphaseResampler(up, down, taps)
    assert r.apply(np.zeros(0, np.complex64)).size == 0
    d_in = fd.DeviceBuffer.from_array(x)
    d_out = fd.DeviceBuffer(r.needed_out_count(x.size))
    n = r.apply_device(d_in, x.size, d_out)
    fd.device_sync()
    want = O.Resampler(up, down, taps).apply(x)
    assert n == want.size
    assert_close_rms(d_out.download(n), want, rms(x))


def test_capacity_too_small_is_an_error():
    r = fd.PolyphaseResampler(1, 2, np.ones(3, np.complex64))
    d_in = fd.DeviceBuffer.from_array(np.ones(100, np.complex64))
    d_out = fd.DeviceBuffer(10)
    with pytest.raises(fd.FdspError):
        r.apply_device(d_in, 100, d_out)
    # the failed call left the state untouched
    assert r.state()[2] == 0


def test_apply_chunks_equals_per_chunk_calls():
    # one launch over 16 consecutive chunks (fdsp_resampler_apply_chunks_device)
    # gives the per-chunk calls' outputs, back to back, and their counts
    w = Resampler1ch()
    x = w.input()
    a = fd.PolyphaseResampler(w.up, w.down, w.taps())
    b = fd.PolyphaseResampler(w.up, w.down, w.taps())
    d_in = fd.DeviceBuffer.from_array(x)
    ob = fd.DeviceBuffer(w.n * w.up // w.down + 16)
    s = fd.Stream()
    for rep in range(2):  # the second pass continues the stream
        got_a = []
        for k in range(w.n // w.chunk):
            seg = fd.DeviceBuffer(w.chunk * w.up // w.down + 2)
            cnt = a.apply_device(d_in, w.chunk, seg, s, in_offset=k * w.chunk)
            s.sync()
            got_a.append(seg.download(cnt))
        counts = b.apply_chunks_device(d_in, w.chunk, w.n // w.chunk, ob, s)
        s.sync()
        assert counts == [g.size for g in got_a]
        np.testing.assert_array_equal(ob.download(sum(counts)), np.concatenate(got_a))
        assert a.state() == b.state()


@pytest.mark.parametrize("up,down,cut", [(4, 5, 70_001), (1, 7, 3), (3, 2, 99_999), (160, 147, 12_345)])
def test_set_state_round_trip(up, down, cut):
    # a handle restored from another's (coeff_phase, xOffset) and its last
    # coeffs_per_phase - 1 inputs continues the stream exactly as the original
    # and as the oracle (fdsp_resampler_set_state)
    taps = kaiser_sinc(24 * up, 0.5 / max(up, down), 8.0, gain=up).astype(np.complex64)
    x = cwgn(43, 200_000)
    a = fd.PolyphaseResampler(up, down, taps)
    ref = O.Resampler(up, down, taps)
    a.apply(x[:cut])
    ref.apply(x[:cut])
    ph, xo, _, _ = a.state()
    cpp = a.coeffs_per_phase()
    # the reference's state: the last cpp - 1 inputs, zeros before the stream's start
    hist = np.concatenate([np.zeros(max(0, cpp - 1 - cut), np.complex64), x[max(0, cut - (cpp - 1)):cut]])
    b = fd.PolyphaseResampler(up, down, taps)
    b.set_state(ph, xo, hist)
    assert b.state()[:2] == (ph, xo)
    for lo, hi in chunks(x.size - cut, [50_000, 1, 17]):
        ya, yb = a.apply(x[cut + lo:cut + hi]), b.apply(x[cut + lo:cut + hi])
        r = ref.apply(x[cut + lo:cut + hi])
        np.testing.assert_array_equal(yb, ya)
        assert yb.size == r.size
        assert_close_rms(yb, r, rms(x))
        assert b.state()[:2] == a.state()[:2] == ref.state_fields


def test_set_state_rejects_bad_fields():
    r = fd.PolyphaseResampler(4, 5, np.ones(96, np.complex64))
    with pytest.raises(fd.FdspError):
        r.set_state(4, 0, np.zeros(23, np.complex64))   # coeff_phase >= up
    with pytest.raises(fd.FdspError):
        r.set_state(0, 0, np.zeros(5, np.complex64))    # history is cpp - 1 = 23 samples
    assert r.state()[:2] == (0, 0)
