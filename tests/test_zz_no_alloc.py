"""fdsp_iir_apply_device_noalloc against fdsp_iir_apply_device, through the C ABI.

The no-allocation form runs the same kernels with their per-segment state
records and block flags in the library's static per-device array, so its
outputs must equal the allocating form's bit for bit (np.testing.
assert_array_equal), chunk after chunk, on two calls in a row with the second
smaller than the first, and the exact plan's outputs must equal the oracle's.
Filters: the FM de-emphasis (first order; verified segments, whose finishing
wave reads the records) and a 4th-order Butterworth lowpass in the automatic
mode (float outputs from unverified segments) and in exact mode (one
sequential wave). A chunk too large for the static array fails with
FDSP_ERANGE and leaves the stream state as it was.
"""
import numpy as np
import pytest

import libfalcon_dsp_amd as fd
from libfalcon_dsp_amd.signals import cwgn
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _filters():
    b0, a1 = fd.fm_deemphasis(75e-6)
    from scipy.signal import butter
    bb, ab = butter(4, 0.2)
    return {
        "deemphasis": ([b0], [1.0, a1], None),
        "butter4_auto": (bb, ab, None),
        "butter4_exact": (bb, ab, True),
    }


@pytest.mark.parametrize("name", ["deemphasis", "butter4_auto", "butter4_exact"])
def test_noalloc_equals_allocating_form(name):
    b, a, exact = _filters()[name]
    sizes = [(1 << 20) + 5, 3001]  # the second call is the smaller one
    x = cwgn(91, sum(sizes))
    f_ref, f_na = fd.IirFilter(b, a), fd.IirFilter(b, a)
    if exact is not None:
        f_ref.set_exact(exact)
        f_na.set_exact(exact)
    oracle = O.Iir(np.asarray(b, np.complex64), np.asarray(a, np.complex64))
    s = fd.Stream()
    i = 0
    for n in sizes:
        print(f"noalloc {name}: chunk of {n}", flush=True)
        d_in = fd.DeviceBuffer.from_array(x[i:i + n])
        d_ref, d_na = fd.DeviceBuffer(n), fd.DeviceBuffer(n)
        f_ref.apply_device(d_in, n, d_ref, s)
        f_na.apply_device_noalloc(d_in, n, d_na, s)
        s.sync()
        got, want = d_na.download(n), d_ref.download(n)
        np.testing.assert_array_equal(got, want)
        r = oracle.apply(x[i:i + n])
        if name != "butter4_auto":
            np.testing.assert_array_equal(got, r)
        i += n


def test_noalloc_rejects_a_chunk_beyond_its_array():
    from scipy.signal import butter
    bb, ab = butter(8, 0.2)  # order 8: 2^20 samples need more than the 16 MiB
    f, g = fd.IirFilter(bb, ab), fd.IirFilter(bb, ab)
    n = 1 << 20
    x = cwgn(92, n + 4000)
    d_in = fd.DeviceBuffer.from_array(x)
    d_out = fd.DeviceBuffer(n)
    s = fd.Stream()
    with pytest.raises(fd.FdspError, match="FDSP_ERANGE"):
        f.apply_device_noalloc(d_in, n, d_out, s)
    # nothing moved: both handles continue alike from the stream's start
    d_a, d_b = fd.DeviceBuffer(4000), fd.DeviceBuffer(4000)
    f.apply_device_noalloc(d_in, 4000, d_a, s)
    g.apply_device(d_in, 4000, d_b, s)
    s.sync()
    np.testing.assert_array_equal(d_a.download(4000), d_b.download(4000))
