"""FM demodulation chain on the GPU vs the oracle (fdsp_iir1_*, fdsp_fm_demod_*).

The oracle equals the compiled reference bit for bit on both (tests/test_fm_cpu.py).
Bounds:
- the first-order IIR (de-emphasis) is bit-identical to the oracle, chunk
  after chunk, including chunks shorter than the kernel's warm-up;
- the demodulator's float audio (before scaling) is within 1e-4 x its RMS of
  the oracle's, with exact counts; its int16 output is within 1 LSB (the
  float chain's rounding differs from the reference's, and the truncation to
  int16 turns a difference across an integer into one step), with at most 5%
  of samples one step apart.
"""
import os

import numpy as np
import pytest

import libfalcon_dsp_amd as fd
from libfalcon_dsp_amd.signals import cwgn, fm_signal
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TAPS = np.load(os.path.join(ROOT, "tests", "golden", "fm_taps.npz"))
FS = 1_140_000


def fm_taps():
    iso = (int(TAPS["iso_rates"][2]), int(TAPS["iso_rates"][3]), TAPS["iso_taps"])
    mono = (int(TAPS["mono_rates"][2]), int(TAPS["mono_rates"][3]), TAPS["mono_taps"])
    return iso, mono


def fm_input(seed, n, offset):
    sig, _ = fm_signal(n, FS, [400.0, 1300.0, 2700.0], 75e3, carrier_hz=offset)
    return (sig + 0.02 * cwgn(seed, n)).astype(np.complex64)


@pytest.mark.parametrize("tau", [75e-6, 50e-6])
def test_iir1_bit_exact_vs_oracle(tau):
    b0, a1 = fd.fm_deemphasis(tau)
    x = (np.random.default_rng(7).standard_normal(300_000) * 0.8).astype(np.float32)
    sizes = [1, 17, 1000, 5000, 150_000, 2, 143_980]
    g, o = fd.Iir1(b0, a1), O.Iir([b0], [1, a1])
    i = 0
    for k, s in enumerate(sizes):
        np.testing.assert_array_equal(g.apply(x[i:i + s]), o.apply_f32(x[i:i + s]), err_msg=f"chunk {k}")
        i += s
    g.reset_state()
    o.reset()
    np.testing.assert_array_equal(g.apply(x[:70_000]), o.apply_f32(x[:70_000]))
    # a slower pole is exact too (longer warm-up, verified merge)
    g2, o2 = fd.Iir1(0.03, -0.97), O.Iir([0.03], [1, -0.97])
    np.testing.assert_array_equal(g2.apply(x[:100_000]), o2.apply_f32(x[:100_000]))


@pytest.mark.parametrize("tau", [75e-6, 50e-6])
def test_fm_demod_vs_oracle(tau):
    iso, mono = fm_taps()
    offset = -200_000
    n = FS
    x = fm_input(9, n, -offset)
    sizes = [262_144, 1, 7, 131_072, 400_000, n - 793_224]
    g, o = fd.FmDemodulator(FS, offset, iso, mono, tau), O.Fm(FS, offset, tau, iso, mono)
    i, total, differ = 0, 0, 0
    for k, s in enumerate(sizes):
        c = x[i:i + s]
        gi, ga = g.demod_mono(c, audio=True)
        oi, oa = o.demod(c)
        assert gi.size == oi.size, f"chunk {k}: {gi.size} vs {oi.size}"
        if oi.size:
            r = np.sqrt(np.mean(oa.astype(np.float64) ** 2))
            assert np.max(np.abs(ga - oa)) <= 1e-4 * r, f"chunk {k}: audio error {np.max(np.abs(ga - oa)):.3e}"
            d = np.abs(gi.astype(np.int32) - oi)
            assert d.max() <= 1, f"chunk {k}: int16 max diff {d.max()}"
            differ += int(np.count_nonzero(d))
            total += d.size
        i += s
    assert total == 45_600  # one second of audio
    assert differ <= 0.05 * total, f"{differ} of {total} int16 samples one step apart"


def test_fm_demod_long_chunks():
    # the rescale's three forms: up to 16 audio samples a thread in registers
    # (the chunks above), one block re-reading the chunk (20,000 audio
    # samples), and the multi-block max then rescale (68,000)
    iso, mono = fm_taps()
    offset = -200_000
    x = fm_input(21, 2_200_000, -offset)
    g, o = fd.FmDemodulator(FS, offset, iso, mono), O.Fm(FS, offset, 75e-6, iso, mono)
    for k, (lo, hi) in enumerate([(0, 500_000), (500_000, 2_200_000)]):
        gi, ga = g.demod_mono(x[lo:hi], audio=True)
        oi, oa = o.demod(x[lo:hi])
        assert gi.size == oi.size, f"chunk {k}: {gi.size} vs {oi.size}"
        assert gi.size > (16_384, 65_536)[k]  # past the register form, then past one block
        r = np.sqrt(np.mean(oa.astype(np.float64) ** 2))
        assert np.max(np.abs(ga - oa)) <= 1e-4 * r, f"chunk {k}: audio error {np.max(np.abs(ga - oa)):.3e}"
        d = np.abs(gi.astype(np.int32) - oi)
        assert d.max() <= 1, f"chunk {k}: int16 max diff {d.max()}"
        assert np.count_nonzero(d) <= 0.05 * d.size


def test_fm_demod_device_and_errors():
    iso, mono = fm_taps()
    x = fm_input(3, 200_000, 50_000)
    g = fd.FmDemodulator(FS, -50_000, iso, mono)
    cap = g.out_count(x.size)
    d_in = fd.DeviceBuffer.from_array(x)
    d_out = fd.DeviceBuffer(cap, np.int16)
    s = fd.Stream()
    import ctypes as C
    n = C.c_size_t()
    fd.lib().fdsp_fm_demod_apply_device(g.h, d_in.ptr, x.size, d_out.ptr, None, cap, C.byref(n), s.h)
    s.sync()
    got = d_out.download(n.value)
    want, _ = O.Fm(FS, -50_000, 75e-6, iso, mono).demod(x)
    assert got.size == want.size and np.max(np.abs(got.astype(np.int32) - want)) <= 1
    with pytest.raises(fd.FdspError):
        g.demod_mono(np.zeros(0, np.complex64))
    # capacity below what the chunk can emit: FDSP_ERANGE before any work, so
    # the next chunk continues the stream as if the call had not happened
    h2, o2 = fd.FmDemodulator(FS, -50_000, iso, mono), O.Fm(FS, -50_000, 75e-6, iso, mono)
    first = x[:100_000]
    small = np.empty(1, np.int16)
    rc = fd.lib().fdsp_fm_demod_apply(h2.h, first.ctypes.data, first.size, small.ctypes.data, None, 1, C.byref(n))
    assert rc == fd._native.FDSP_ERANGE
    a, _ = o2.demod(first)
    assert np.max(np.abs(h2.demod_mono(first).astype(np.int32) - a)) <= 1


def test_fm_default_filters_are_the_predefined_ones():
    """FmDemodulator without filters takes the predefined 64-tap REMEZ pair,
    as the reference's initialize does; the result equals the fixture's."""
    iso, mono = fm_taps()
    x = fm_input(12, 200_000, 150_000)
    a = fd.FmDemodulator(FS, -150_000)
    b = fd.FmDemodulator(FS, -150_000, iso, mono)
    for part in (slice(0, 80_000), slice(80_000, 200_000)):
        np.testing.assert_array_equal(a.demod_mono(x[part]), b.demod_mono(x[part]))

