"""FIR filter, FIR bank and polar discriminator on the GPU
(reference src/transform/falcon_dsp_fir_filter.cc:177-284,
src/transform/falcon_dsp_polar_discriminator.cc:134-171; KATs from
test/tests/transform/falcon_dsp_fir_filter_cpp_unit_tests.cc:147-306 and
falcon_dsp_polar_discriminator_cpp_unit_tests.cc:78-138)."""
import numpy as np
import pytest

import libfalcon_dsp_amd as fd
from libfalcon_dsp_amd.signals import cwgn, kaiser_sinc
from oracle import oracle as O
from tests.helpers import assert_close_rms, chunks, rms

pytestmark = pytest.mark.gpu

H = np.asarray([0.5, 2.0, 4.0], np.complex64)


def test_fir_kats_exact():
    f = fd.FirFilter(H)
    np.testing.assert_array_equal(f.apply(np.ones(3, np.complex64)).real, [0.5, 2.5, 6.5])
    f = fd.FirFilter(H)
    np.testing.assert_array_equal(f.apply(np.ones(8, np.complex64)).real, [0.5, 2.5] + [6.5] * 6)
    x = np.asarray([1, 1, 1, 2, 2, 2, 5, 5, 5, 1, 1], np.complex64)
    np.testing.assert_array_equal(fd.FirFilter(H).apply(x).real,
                                  [0.5, 2.5, 6.5, 7.0, 9.0, 13.0, 14.5, 20.5, 32.5, 30.5, 22.5])


def test_fir_one_sample_at_a_time():
    f = fd.FirFilter(H)
    got = [f.apply(np.ones(1, np.complex64))[0].real for _ in range(8)]
    assert got == [0.5, 2.5] + [6.5] * 6


@pytest.mark.parametrize("m", [1, 2, 31, 256, 1000])
def test_fir_complex_taps_chunked(m):
    rng = np.random.default_rng(m)
    h = (rng.standard_normal(m) + 1j * rng.standard_normal(m)).astype(np.complex64) / np.sqrt(m)
    x = cwgn(30 + m, 20_000)
    gpu, ref = fd.FirFilter(h), O.Fir(h)
    xr = rms(x) * np.sqrt(np.sum(np.abs(h) ** 2))  # output scale
    for lo, hi in chunks(x.size, [1, 5, 3000, m + 3, 7]):
        assert_close_rms(gpu.apply(x[lo:hi]), ref.apply(x[lo:hi]), max(xr, 1e-30), what=f"[{lo},{hi})")


def test_fir_empty_is_an_error():
    with pytest.raises(fd.FdspError):
        fd.FirFilter(H).apply(np.zeros(0, np.complex64))


def test_fir_bank_against_oracle():
    n_chan, m, n = 6, 512, 50_000
    rng = np.random.default_rng(7)
    taps = np.stack([kaiser_sinc(m, 0.05 + 0.4 * rng.random(), 8.0) for _ in range(n_chan)])
    x = np.stack([cwgn(3, n, offset=c * n) for c in range(n_chan)])
    bank = fd.FirBank(taps)
    refs = [O.Fir(taps[c].astype(np.complex64)) for c in range(n_chan)]
    for lo, hi in ((0, 20_000), (20_000, 20_001), (20_001, n)):
        seg = np.ascontiguousarray(x[:, lo:hi])
        y = bank.apply(seg)
        for c in range(n_chan):
            assert_close_rms(y[c], refs[c].apply(seg[c]), rms(x[c]), what=f"chan {c} [{lo},{hi})")


@pytest.mark.parametrize("m", [1, 7, 100, 1000, 3000])
def test_fir_bank_tap_counts_and_strides(m):
    """Tap counts off the matrix-core K step of 4 (1 to 1000 taps run on
    FP32 matrix cores), 3000 taps (a window past the LDS budget: the VALU
    form), chunks shorter than a 4096-output tile, channel strides above the
    chunk length; every channel against the oracle at 1e-4 x RMS."""
    n_chan, n = 3, 9000
    rng = np.random.default_rng(m)
    taps = np.stack([kaiser_sinc(m, 0.05 + 0.4 * rng.random(), 8.0) if m > 1 else np.ones(1) for _ in range(n_chan)])
    x = np.stack([cwgn(40 + c, n) for c in range(n_chan)])
    bank = fd.FirBank(taps)
    refs = [O.Fir(taps[c].astype(np.complex64)) for c in range(n_chan)]
    stride = 5000
    for lo, hi in ((0, 300), (300, 4396), (4396, 9000)):
        k = hi - lo
        d_in, d_out = fd.DeviceBuffer(n_chan * stride), fd.DeviceBuffer(n_chan * stride)
        for c in range(n_chan):
            d_in.upload(x[c, lo:hi], offset=c * stride)
        bank.apply_device(d_in, k, stride, d_out, stride)
        fd.device_sync()
        for c in range(n_chan):
            got = d_out.download(k, offset=c * stride)
            want = refs[c].apply(x[c, lo:hi])
            scale = rms(x[c]) * np.sqrt(np.sum(taps[c] ** 2))
            assert_close_rms(got, want, scale, what=f"m={m} chan {c} [{lo},{hi})")


def test_discriminator_kat_and_streaming():
    c = np.asarray([88 - 85j, 22 + 97j, 54 + 91j, 6 - 13j, -29 + 58j], np.complex64)
    want = [2.11582422, -0.31252633, -2.17362758, -3.11035282]
    d = fd.PolarDiscriminator()
    np.testing.assert_allclose(d.apply(c), want, atol=1e-4)
    # later calls pair the first sample with the saved last one
    x = cwgn(12, 10_000)
    g, r = fd.PolarDiscriminator(), O.Discriminator()
    for lo, hi in chunks(x.size, [1, 2, 999, 4096]):
        a, b = g.apply(x[lo:hi]), r.apply(x[lo:hi])
        assert a.size == b.size
        # compare angles modulo 2 pi: a product landing on the negative real
        # axis may round to +pi on one side and -pi on the other
        d = np.angle(np.exp(1j * (a.astype(np.float64) - b)))
        assert d.size == 0 or np.max(np.abs(d)) <= 1e-5
    g.reset_state()
    assert g.apply(x[:10]).size == 9
