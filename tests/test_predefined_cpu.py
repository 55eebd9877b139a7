"""Predefined resampling filters (get_resample_fir_params) on the CPU.

The table is designed at build time by libfalcon_dsp_amd/predefined.py from
the reference's documented procedure. Pinned by:
- the compiled reference (oracle/_ref): every key the reference's table holds
  exists here with the same up/down rates and bit-identical taps, and no key
  exists here that the reference lacks;
- the designs' own properties, independent of the procedure's text: the long
  least-squares filters have DC gain equal to the up rate (the procedure
  scales by it, so a sum of taps off by a factor shows), the equiripple
  filters pass DC with gain near 1 and stop below -20 dB past the transition;
- the FM fixture (tests/golden/fm_taps.npz, written from oracle/_ref) equals
  the two 64-tap REMEZ filters served here.
The lookup is host code: these tests call the library without a GPU.
"""
import os

import numpy as np
import pytest

import libfalcon_dsp_amd as fd
from libfalcon_dsp_amd import predefined as P
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")
TAPS_KEYS = (-1, 16, 32, 64, 128, 256)


def served(fin, fout, taps, source):
    try:
        return fd.resample_fir_params(fin, fout, taps, source)
    except fd.FdspError:
        return None


@needs_ref
def test_table_equals_the_reference_table():
    n = 0
    for fin, fout in P.RATE_PAIRS + [(1_000_000, 250_000), (48_000, 8_000)]:
        for taps in TAPS_KEYS:
            for source in (fd.FILTER_REMEZ, fd.FILTER_FIRLS):
                mine, ref = served(fin, fout, taps, source), O.ref_predefined_filter(fin, fout, taps, source)
                assert (mine is None) == (ref is None), (fin, fout, taps, source)
                if ref is None:
                    continue
                assert mine[:2] == ref[:2], (fin, fout, taps, source)
                np.testing.assert_array_equal(mine[2], ref[2], err_msg=str((fin, fout, taps, source)))
                n += 1
    assert n == 51


def test_design_properties():
    for fin, fout in P.RATE_PAIRS:
        up, down, h = fd.resample_fir_params(fin, fout, 64, fd.FILTER_FIRLS)
        assert up * fin == down * fout and np.gcd(up, down) == 1
        assert h.size == 65 and not h.imag.any()
        # the long least-squares design (20 max(up, down) + 1 taps): DC gain = the up rate
        up, down, h = fd.resample_fir_params(fin, fout, fd.FILTER_TAPS_OPTIMAL, fd.FILTER_FIRLS)
        assert h.size == 20 * max(up, down) + 1
        assert abs(h.real.sum() - up) < 0.02 * up
        up, down, h = fd.resample_fir_params(fin, fout, 64, fd.FILTER_REMEZ)
        w = np.fft.rfft(h.real.astype(np.float64), 1 << 16)
        f = np.arange(w.size) / (1 << 16) * fin
        assert abs(abs(w[0]) - 1.0) < 0.05
        stop = np.abs(w[f >= fout + (fin / 2 - fout) / 4])
        assert 20 * np.log10(stop.max()) < -20  # stopband below -20 dB
    with pytest.raises(fd.FdspError):
        fd.resample_fir_params(1_140_000, 228_000, 256, fd.FILTER_REMEZ)  # designed only up to 200 taps
    with pytest.raises(fd.FdspError):
        fd.resample_fir_params(48_000, 8_000)


def test_fm_fixture_is_the_served_filters():
    t = np.load(os.path.join(ROOT, "tests", "golden", "fm_taps.npz"))
    for name, (fin, fout) in (("iso", (1_140_000, 228_000)), ("mono", (228_000, 45_600))):
        up, down, h = fd.resample_fir_params(fin, fout)
        assert (up, down) == (int(t[f"{name}_rates"][2]), int(t[f"{name}_rates"][3]))
        np.testing.assert_array_equal(h, t[f"{name}_taps"])
