"""IIR bank (fdsp_iir_bank_*): many streams, each against its own oracle filter.

Bound: every stream's outputs equal the oracle's sequential loop on that
stream exactly (np.testing.assert_array_equal, so +0 and -0 compare equal),
chunk after chunk, for the filters whose single-stream GPU form is only
exact as one sequential wave (4th-order Butterworth) and for random filters
of order 0 to 16 with real and complex coefficients and a[0] != 1; stream
counts that leave a partial block of 64; chunks of 1 sample up to several
tiles; complex float and complex int16 samples; device rows with strides
above n; and reset.
"""
import numpy as np
import pytest

import libfalcon_dsp_amd as fd
from libfalcon_dsp_amd.signals import cwgn
from oracle import oracle as O
from tests.test_gpu_iir import chunks, random_filter

pytestmark = pytest.mark.gpu

SIZES = [1, 15, 16, 17, 1000, 33, 2]


def _oracles(b, a, n_streams):
    return [O.Iir(b, a) for _ in range(n_streams)]


@pytest.mark.parametrize("nb,nfb,cplx,a0,streams", [(5, 4, False, 1.0, 70), (1, 1, False, 1.0, 64),
                                                    (3, 2, False, 2.0, 1), (2, 3, True, 0.7 - 0.2j, 129),
                                                    (4, 13, True, -1.5, 65), (17, 16, False, 1.0, 40),
                                                    (6, 0, True, 1.0, 3)])
def test_bank_equals_oracle_per_stream(nb, nfb, cplx, a0, streams):
    u = np.random.default_rng(nb * 1000 + nfb * 10 + streams)
    b, a = random_filter(u, nb, nfb, cplx, a0)
    if (nb, nfb) == (5, 4):
        from libfalcon_dsp_amd.workloads import IirOrder4
        b, a = IirOrder4().coeffs()
    print(f"iir bank nb={nb} order={nfb} complex={cplx} streams={streams}", flush=True)
    n = sum(SIZES)
    x = np.stack([cwgn(7 + s, n) for s in range(streams)])
    bank, refs = fd.IirBank(b, a, streams), _oracles(b, a, streams)
    lo = 0
    for k, m in enumerate(SIZES):
        got = bank.apply(x[:, lo:lo + m])
        for s in range(streams):
            np.testing.assert_array_equal(got[s], refs[s].apply(x[s, lo:lo + m]), err_msg=f"chunk {k} stream {s}")
        lo += m


def test_bank_ci16_and_reset():
    from libfalcon_dsp_amd.workloads import IirOrder4
    b, a = IirOrder4().coeffs()
    streams, n = 100, 3000
    x = np.stack([cwgn(50 + s, n) for s in range(streams)])
    xi = np.round(np.stack([x.real, x.imag], -1) * 3000).astype(np.int16)
    bank, refs = fd.IirBank(b, a, streams), _oracles(b, a, streams)
    for k, (lo, hi) in enumerate([(0, 700), (700, 701), (701, 3000)]):
        got = bank.apply_ci16(xi[:, lo:hi])
        for s in range(streams):
            np.testing.assert_array_equal(got[s], refs[s].apply_ci16(xi[s, lo:hi]), err_msg=f"chunk {k} stream {s}")
    bank.reset_state()
    fresh = _oracles(b, a, streams)
    got = bank.apply_ci16(xi[:, :500])
    for s in range(streams):
        np.testing.assert_array_equal(got[s], fresh[s].apply_ci16(xi[s, :500]), err_msg=f"after reset, stream {s}")


def test_bank_device_rows_with_strides():
    # rows of n samples inside longer rows; samples past n are not touched
    from libfalcon_dsp_amd.workloads import IirOrder4
    b, a = IirOrder4().coeffs()
    streams, n, si, so = 66, 1234, 1300, 1250
    x = np.stack([cwgn(80 + s, si) for s in range(streams)])
    d_in = fd.DeviceBuffer.from_array(x.reshape(-1))
    d_out = fd.DeviceBuffer(streams * so)
    fill = np.full(streams * so, 7 + 7j, np.complex64)
    d_out.upload(fill)
    bank, refs = fd.IirBank(b, a, streams), _oracles(b, a, streams)
    s_ = fd.Stream()
    bank.apply_device(d_in, n, si, d_out, so, s_)
    s_.sync()
    got = d_out.download().reshape(streams, so)
    for s in range(streams):
        np.testing.assert_array_equal(got[s, :n], refs[s].apply(x[s, :n]), err_msg=f"stream {s}")
        np.testing.assert_array_equal(got[s, n:], fill[:so - n])
    with pytest.raises(fd.FdspError):
        bank.apply_device(d_in, n, n - 1, d_out, so, s_)


def test_bank_long_rows_match_single_stream_exact_plan():
    # 2^16 samples of 64 streams of the 4th-order Butterworth: each row equals
    # the single-stream exact (sequential) plan
    from libfalcon_dsp_amd.workloads import IirOrder4
    b, a = IirOrder4().coeffs()
    streams, n = 64, 1 << 16
    x = np.stack([cwgn(300 + s, n) for s in range(streams)])
    got = fd.IirBank(b, a, streams).apply(x)
    for s in (0, 1, 31, 63):
        g = fd.IirFilter(b, a)
        g.set_exact(True)
        np.testing.assert_array_equal(got[s], g.apply(x[s]), err_msg=f"stream {s}")
        np.testing.assert_array_equal(got[s], O.Iir(b, a).apply(x[s]), err_msg=f"stream {s} vs oracle")


def test_bank_rejects_bad_arguments():
    with pytest.raises(fd.FdspError):
        fd.IirBank(np.ones(18), [1.0], 4)  # more than 17 b taps
    with pytest.raises(fd.FdspError):
        fd.IirBank([1.0], [1.0], 0)
    with pytest.raises(fd.FdspError):
        fd.IirBank([1.0], [0.0, 1.0], 4)  # a[0] == 0
