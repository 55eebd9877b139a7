"""General IIR filter on the GPU vs the oracle (fdsp_iir_*).

The oracle equals the compiled reference bit for bit on all four sample types
(tests/test_iir_cpu.py). Bound, in exact mode: the GPU outputs equal the
oracle's exactly (np.testing.assert_array_equal, so +0 and -0 compare equal),
chunk after chunk, for stable, slowly decaying and unstable filters, whichever
plan runs them (verified segments for filters whose segments merge, one
sequential wave for the others):
- the reference's known-answer tests;
- random filters of order 0 to 16 with real and complex coefficients, on
  cf32, f32, ci16 and int16 samples, in chunks from 1 sample to 2^18 + 3;
- a pole at radius 0.9995, whose decay outlasts the longest warm-up (it runs
  sequentially);
- a long run of zeros after noise, where the true output decays through
  denormals while a zero-started segment stays at zero (repair path again);
- 2^24 samples of an order-4 filter in one call (the benchmark's size).
In the default automatic mode int16 outputs are exact, and so are float
outputs of filters whose segments merge; other float outputs, and all outputs
in fast mode (set_exact(False), segments unverified), are within 1e-4 x the
output RMS of the oracle's (fast-mode int16 outputs within one step).
"""
import numpy as np
import pytest

import libfalcon_dsp_amd as fd
from libfalcon_dsp_amd.signals import cwgn
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 63, 64, 65, 4095, 4097, 70_000, (1 << 18) + 3]


def chunks(x, sizes):
    i = 0
    for s in sizes:
        yield x[i:i + s]
        i += s


def random_filter(u, nb, nfb, complex_coeffs, a0=1.0, rmax=0.95):
    b = u.standard_normal(nb) * 0.5
    if complex_coeffs:
        b = b + 1j * u.standard_normal(nb) * 0.5
    r = u.uniform(0.2, rmax, nfb)
    poles = r * np.exp(1j * u.uniform(-np.pi, np.pi, nfb))
    if not complex_coeffs:
        half = poles[: nfb // 2]
        poles = np.concatenate([half, np.conj(half)] + ([r[-1:]] if nfb % 2 else []))
    a = np.poly(poles) if nfb else np.array([1.0])
    a = (a if complex_coeffs else a.real) * a0
    return np.asarray(b, np.complex64), np.asarray(a, np.complex64)


def test_reference_known_answers():
    f = fd.IirFilter([1.0], [0.5, 1.0])
    got = f.apply_ci16(np.array([[0, 0], [1, 0], [1, 0], [1, 0], [1, 0]], np.int16))
    np.testing.assert_array_equal(got, [[0, 0], [2, 0], [-2, 0], [6, 0], [-10, 0]])
    got = fd.IirFilter([1.0], [0.5, 1.0]).apply(np.array([0.25, 1, 1, 1, 1], np.complex64))
    np.testing.assert_allclose(got, [0.5, 1.0, 0.0, 2.0, -2.0], atol=1e-4)
    got = fd.IirFilter([1.0], [1.0, 0.9, 0.8]).apply(np.array([0, 1, 1, 1, 1], np.complex64))
    np.testing.assert_allclose(got, [0.0, 1.0, 0.1, 0.11, 0.821], atol=1e-4)
    got = fd.iir_filter([0.3, 0.16, 0.21], [1.0], np.array([0, 10, 15, 11, 13], np.complex64))
    np.testing.assert_allclose(got, [0.0, 3.0, 6.1, 7.8, 8.81], atol=1e-4)
    # the same, exactly as the oracle computes them; and an unstable filter
    # (pole at -2) over 100 samples, still finite
    x = np.ones(100, np.complex64)
    np.testing.assert_array_equal(fd.IirFilter([1.0], [0.5, 1.0]).apply(x), O.Iir([1.0], [0.5, 1.0]).apply(x))


@pytest.mark.parametrize("nb,nfb,cplx,a0", [(1, 1, False, 1.0), (3, 2, False, 2.0), (2, 3, True, 0.7 - 0.2j),
                                            (5, 5, True, 1.0), (1, 8, False, 1.0), (4, 13, True, -1.5),
                                            (8, 16, False, 1.0), (6, 0, True, 1.0)])
def test_random_filters_all_types_vs_oracle(nb, nfb, cplx, a0):
    u = np.random.default_rng(nb * 100 + nfb)
    b, a = random_filter(u, nb, nfb, cplx, a0)
    n = sum(SIZES)
    x = cwgn(nfb + 3, n)
    print(f"iir nb={nb} order={nfb} complex={cplx} a0={a0} n={n}", flush=True)

    def exact():
        g = fd.IirFilter(b, a)
        g.set_exact(True)
        return g, O.Iir(b, a)

    g, o = exact()
    for k, c in enumerate(chunks(x, SIZES)):
        np.testing.assert_array_equal(g.apply(c), o.apply(c), err_msg=f"cf32 chunk {k}")
    xr = x.real.copy()
    g, o = exact()
    for k, c in enumerate(chunks(xr, SIZES[:7])):
        np.testing.assert_array_equal(g.apply_f32(c), o.apply_f32(c), err_msg=f"f32 chunk {k}")
    xi = np.round(np.stack([x.real, x.imag], 1) * 200).astype(np.int16)
    g, o = exact()
    for k, c in enumerate(chunks(xi, SIZES[:8])):
        np.testing.assert_array_equal(g.apply_ci16(c), o.apply_ci16(c), err_msg=f"ci16 chunk {k}")
    x16 = xi[:, 0].copy()
    g, o = fd.IirFilter(b, a), O.Iir(b, a)  # int16 outputs are exact in the default mode
    for k, c in enumerate(chunks(x16, SIZES[:8])):
        np.testing.assert_array_equal(g.apply_i16(c), o.apply_i16(c), err_msg=f"i16 chunk {k}")
    # default mode, float outputs: exact when the segments merge, else within rounding
    g, o = fd.IirFilter(b, a), O.Iir(b, a)
    for k, c in enumerate(chunks(x, SIZES[:8])):
        got, want = g.apply(c), o.apply(c)
        if g.plan == "segments":
            np.testing.assert_array_equal(got, want, err_msg=f"auto cf32 chunk {k}")
        else:
            assert np.max(np.abs(got - want)) <= 1e-4 * np.sqrt(np.mean(np.abs(want) ** 2)), f"auto chunk {k}"


def test_plans():
    b0, a1 = fd.fm_deemphasis(75e-6)
    assert fd.IirFilter([b0], [1, a1]).plan == "segments"
    from libfalcon_dsp_amd.workloads import IirOrder4
    g = fd.IirFilter(*IirOrder4().coeffs())
    assert g.plan == "fast"  # its feedback keeps zero-started segments a few ulp off
    g.set_exact(True)
    assert g.plan == "serial"
    g.set_exact(False)
    assert g.plan == "fast"
    g.set_exact(None)
    assert g.plan == "fast"


def test_slow_pole_is_exact():
    b, a = [0.0005], [1.0, -0.9995]
    g, o = fd.IirFilter(b, a), O.Iir(b, a)
    assert g.warm == 2048 and g.plan == "serial"  # its decay to 2^-44 needs far longer
    x = cwgn(11, 30_000)
    for k, c in enumerate(chunks(x, [5000, 25_000])):
        np.testing.assert_array_equal(g.apply(c), o.apply(c), err_msg=f"chunk {k}")
    # complex pole near the unit circle
    p = 0.9993 * np.exp(0.3j)
    b2, a2 = [1.0], [1.0, -p]
    np.testing.assert_array_equal(fd.IirFilter(b2, a2).apply(x[:20_000]), O.Iir(b2, a2).apply(x[:20_000]))


def test_silence_after_noise_is_exact():
    b0, a1 = fd.fm_deemphasis(75e-6)
    u = np.random.default_rng(5)
    x = np.concatenate([u.standard_normal(10_000), np.zeros(40_000), u.standard_normal(10_000)]).astype(np.float32)
    g, o = fd.IirFilter([b0], [1, a1]), O.Iir([b0], [1, a1])
    for k, c in enumerate(chunks(x, [30_000, 30_000])):
        got, want = g.apply_f32(c), o.apply_f32(c)
        np.testing.assert_array_equal(got, want, err_msg=f"chunk {k}")
    assert np.any((want != 0) & (np.abs(want) < 1e-30)) or np.any((got != 0) & (np.abs(got) < 1e-30))


def test_mixed_types_and_reset():
    u = np.random.default_rng(9)
    b, a = random_filter(u, 3, 4, True)
    x = cwgn(9, 20_000)
    xr = x.real.copy()
    g, o = fd.IirFilter(b, a), O.Iir(b, a)
    g.set_exact(True)
    np.testing.assert_array_equal(g.apply(x[:5000]), o.apply(x[:5000]))
    np.testing.assert_array_equal(g.apply_f32(xr[5000:12_000]), o.apply_f32(xr[5000:12_000]))
    np.testing.assert_array_equal(g.apply(x[12_000:]), o.apply(x[12_000:]))
    g.reset_state()
    o.reset()
    np.testing.assert_array_equal(g.apply(x[:9000]), o.apply(x[:9000]))


@pytest.mark.parametrize("seed", [0, 1])
def test_fast_mode_within_float_rounding(seed):
    from libfalcon_dsp_amd.workloads import IirOrder4
    if seed == 0:
        b, a = IirOrder4().coeffs()
    else:
        b, a = random_filter(np.random.default_rng(77), 4, 6, True)
    x = cwgn(30 + seed, 300_000)
    g, o = fd.IirFilter(b, a), O.Iir(b, a)
    g.set_exact(False)
    for k, c in enumerate(chunks(x, [1, 100, 4097, 95_802, 200_000])):
        got, want = g.apply(c), o.apply(c)
        rms = np.sqrt(np.mean(np.abs(want) ** 2))
        assert np.max(np.abs(got - want)) <= 1e-4 * rms, f"chunk {k}"
    xi = np.round(np.stack([x.real, x.imag], 1) * 300).astype(np.int16)
    g, o = fd.IirFilter(b, a), O.Iir(b, a)
    g.set_exact(False)
    d = np.abs(g.apply_ci16(xi).astype(np.int32) - o.apply_ci16(xi).astype(np.int32))
    assert d.max() <= 1


def test_device_form_full_size():
    u = np.random.default_rng(4)
    b, a = random_filter(u, 5, 4, False)
    n = 1 << 24
    x = cwgn(21, n)
    want = O.Iir(b, a).apply(x)
    d_in, d_out = fd.DeviceBuffer(n), fd.DeviceBuffer(n)
    d_in.upload(x)
    g = fd.IirFilter(b, a)
    g.apply_device(d_in, n, d_out)  # default mode: float rounding for this filter
    fd.device_sync()
    got = d_out.download(n)
    assert np.max(np.abs(got - want)) <= 1e-4 * np.sqrt(np.mean(np.abs(want) ** 2))
    g = fd.IirFilter(b, a)
    g.set_exact(True)
    g.apply_device(d_in, n, d_out)
    fd.device_sync()
    np.testing.assert_array_equal(d_out.download(n), want)


def test_errors():
    with pytest.raises(fd.FdspError):
        fd.IirFilter([1.0], [0.0, 1.0])  # a[0] = 0
    with pytest.raises(fd.FdspError):
        fd.IirFilter(np.ones(257), [1.0])
    with pytest.raises(fd.FdspError):
        fd.IirFilter([1.0], np.ones(18))
    g = fd.IirFilter([1.0], [1.0, -0.5])
    with pytest.raises(fd.FdspError):
        g.apply(np.zeros(0, np.complex64))


def test_reset_after_apply_on_a_caller_stream():
    # reset waits for a launch still running on the caller's stream (which
    # writes the carried state) before it clears that state
    from scipy.signal import butter
    b, a = butter(4, 0.2)
    x = cwgn(95, 1 << 22)
    f, g = fd.IirFilter(b, a), fd.IirFilter(b, a)
    d_in = fd.DeviceBuffer.from_array(x)
    d_out = fd.DeviceBuffer(x.size)
    s = fd.Stream()
    f.apply_device(d_in, x.size, d_out, s)  # no wait before the reset
    f.reset_state()
    ya, yb = fd.DeviceBuffer(5000), fd.DeviceBuffer(5000)
    f.apply_device(d_in, 5000, ya, s)
    g.apply_device(d_in, 5000, yb, s)
    s.sync()
    np.testing.assert_array_equal(ya.download(5000), yb.download(5000))
