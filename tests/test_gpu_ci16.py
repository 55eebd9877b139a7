"""Complex int16 entry points on the GPU vs the oracle (fdsp_*_apply_ci16).

The reference's int16 overloads convert to float, run the float path and
truncate the outputs toward zero (src/transform/falcon_dsp_freq_shift.cc:133-162,
falcon_dsp_fir_filter.cc:151-175, falcon_dsp_polar_discriminator.cc:117-132).
Bounds:
- int16 outputs of the shift and the FIR equal the oracle's bit for bit: the
  int16 kernels compute as the reference does (FP64 angle and sincos rounded
  to float per sample, unfused float complex products, taps summed in order),
  and the oracle equals the compiled reference on these paths;
- the same outputs whatever the chunking, and at saturation;
- discriminator angles within 1e-5 rad;
- stream state (shift counter, FIR history, carried sample) shared with the
  cf32 calls on the same handle, chunk after chunk.
"""
import numpy as np
import pytest

import libfalcon_dsp_amd as fd
from oracle import oracle as O
from libfalcon_dsp_amd.signals import cwgn

pytestmark = pytest.mark.gpu


def ci16_noise(seed, n, scale):
    x = cwgn(seed, n)
    return np.stack([np.round(x.real * scale), np.round(x.imag * scale)], axis=1).clip(-32768, 32767).astype(np.int16)


def as_cf(x):
    return (x[:, 0].astype(np.float32) + 1j * x[:, 1].astype(np.float32)).astype(np.complex64)


def same_int(a, b, what):
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    d = np.abs(a.astype(np.int32) - b.astype(np.int32))
    assert np.count_nonzero(d) == 0, f"{what}: {np.count_nonzero(d)} of {d.size} differ (max {d.max()})"


SIZES = [65536, 1, 7, 100003, 4096, 2]


def test_freq_shift_ci16_chunked_vs_oracle():
    fs = 25_000_000
    x = ci16_noise(1, sum(SIZES), 6000.0)
    # R = 202500, 250000 and 3571429 take the host-built phasor table; 5 Hz
    # (R = 5e6, above 2^22) takes the FP64 sincos on the device
    for f in (1_234_567, -9_999_999, 7, 5):
        g, o = fd.FreqShift(fs, f), O.FreqShift(fs, f)
        got, want, i = [], [], 0
        for s in SIZES:
            got.append(g.apply_ci16(x[i:i + s]))
            want.append(o.apply_ci16(x[i:i + s]))
            i += s
        same_int(np.concatenate(got), np.concatenate(want), f"shift {f}")
        assert g.samples_handled() == int(o.samples_handled)
        same_int(fd.FreqShift(fs, f).apply_ci16(x), np.concatenate(got), f"shift {f}: one call vs chunks")


def test_multi_shift_ci16_equals_single_shifts():
    fs, shifts = 25_000_000, [-10_000_000, -3_333_333, 1, 6_666_667]
    x = ci16_noise(2, 70_001, 9000.0)
    multi = fd.FreqShift(fs, shifts).apply_ci16(x)
    for f, y in zip(shifts, multi):
        same_int(y, O.FreqShift(fs, f).apply_ci16(x), f"shift {f}")


def test_fir_ci16_exact_and_vs_oracle():
    h = np.array([0.5, 2.0, 4.0], dtype=np.complex64)
    x = np.array([[1, -1]] * 3 + [[2, -2]] * 3 + [[5, -5]] * 3 + [[1, -1]] * 2, dtype=np.int16)
    want = np.array([0, 2, 6, 7, 9, 13, 14, 20, 32, 30, 22], dtype=np.int16)
    np.testing.assert_array_equal(fd.FirFilter(h).apply_ci16(x), np.stack([want, -want], axis=1))
    # saturation
    sat = fd.FirFilter(np.array([4.0], dtype=np.complex64)).apply_ci16(np.array([[20000, -20000]] * 5, dtype=np.int16))
    assert (sat == [32767, -32768]).all()
    # random complex taps, chunks, history shared with cf32 calls on the same handle
    u = np.random.default_rng(3)
    m = 37
    taps = ((u.standard_normal(m) + 1j * u.standard_normal(m)) / np.sqrt(2 * m)).astype(np.complex64)
    x = ci16_noise(4, sum(SIZES), 500.0)
    g, o = fd.FirFilter(taps), O.Fir(taps)
    i = 0
    for k, s in enumerate(SIZES):
        c = x[i:i + s]
        if k == 3:  # a float chunk in the middle: the history carries across
            y, w = g.apply(as_cf(c)), o.apply(as_cf(c))
            assert np.max(np.abs(y - w)) <= 1e-4 * np.sqrt(np.mean(np.abs(w) ** 2) + 1)
        else:
            same_int(g.apply_ci16(c), o.apply_ci16(c), f"fir chunk {k}")
        i += s


def test_discriminator_ci16_vs_oracle():
    x = ci16_noise(5, sum(SIZES), 3000.0)
    g, o = fd.PolarDiscriminator(), O.Discriminator()
    i = 0
    for k, s in enumerate(SIZES):
        c = x[i:i + s]
        if k == 2:
            a, b = g.apply(as_cf(c)), o.apply(as_cf(c))
        else:
            a, b = g.apply_ci16(c), o.apply_ci16(c)
        assert a.size == b.size == (s - 1 if k == 0 else s)
        if a.size:
            assert np.max(np.abs(np.angle(np.exp(1j * (a.astype(np.float64) - b))))) <= 1e-5, f"chunk {k}"
        i += s


def test_ci16_device_entry_points():
    n = 50_000
    x = ci16_noise(6, n, 7000.0)
    d_in = fd.DeviceBuffer.from_array(x.view(fd.CI16).ravel())
    assert d_in.count == n
    d_out = fd.DeviceBuffer(2 * n, fd.CI16)
    s = fd.Stream()
    shifts = [1_000_000, -2_000_000]
    fd.FreqShift(25_000_000, shifts).apply_ci16_device(d_in, n, d_out, n, s)
    s.sync()
    got = d_out.download().view(np.int16).reshape(2, n, 2)
    for c, f in enumerate(shifts):
        same_int(got[c], O.FreqShift(25_000_000, f).apply_ci16(x), f"device shift {f}")
    taps = np.array([0.25, -0.5j, 0.125], dtype=np.complex64)
    d_f = fd.DeviceBuffer(n, fd.CI16)
    fd.FirFilter(taps).apply_ci16_device(d_in, n, d_f, s)
    s.sync()
    same_int(d_f.download().view(np.int16).reshape(n, 2), O.Fir(taps).apply_ci16(x), "device fir")


def test_ci16_argument_errors():
    g = fd.FreqShift(25_000_000, 1000)
    with pytest.raises(fd.FdspError):
        g.apply_ci16(np.zeros((0, 2), dtype=np.int16))
    with pytest.raises(ValueError):
        g.apply_ci16(np.zeros(8, dtype=np.int16))
