"""FM demodulation chain on the CPU: the oracle's IIR and FM demodulator.

Pinned by:
- the compiled reference (oracle/_ref): falcon_dsp_iir_filter (float and
  complex overloads, chunked, a[0] normalisation) and
  falcon_dsp_fm_demodulator::demod_mono on synthetic FM, chunk after chunk,
  equal bit for bit;
- scipy.signal.lfilter in FP64 for the IIR recurrence;
- the de-emphasis coefficients the reference writes as float literals
  (src/signals/falcon_dsp_fm_demod.cc:152-165), reproduced from
  x = exp(-1 / (200000 tau));
- tests/golden/fm_taps.npz, the reference's predefined REMEZ-64 filters read
  through oracle/_ref (tests/golden/make_fm_taps.py).
"""
import os

import numpy as np
import pytest
from scipy import signal as sps

import libfalcon_dsp_amd as fd
from libfalcon_dsp_amd.signals import cwgn, fm_signal
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TAPS = np.load(os.path.join(ROOT, "tests", "golden", "fm_taps.npz"))
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")
FS = 1_140_000


def fm_taps():
    iso = (int(TAPS["iso_rates"][2]), int(TAPS["iso_rates"][3]), TAPS["iso_taps"])
    mono = (int(TAPS["mono_rates"][2]), int(TAPS["mono_rates"][3]), TAPS["mono_taps"])
    return iso, mono


def fm_input(seed, n, offset):
    sig, _ = fm_signal(n, FS, [400.0, 1300.0, 2700.0], 75e3, carrier_hz=offset)
    return (sig + 0.02 * cwgn(seed, n)).astype(np.complex64)


def chunked(f, x, sizes):
    out, i = [], 0
    for s in sizes:
        out.append(f(x[i:i + s]))
        i += s
    return out


def test_deemphasis_coefficients_equal_the_reference_literals():
    want = {75e-6: (np.float32(0.064493014968382223), np.float32(-0.93550698503161778)),
            50e-6: (np.float32(0.095162581964040482), np.float32(-0.90483741803595952))}
    for tau, (b0, a1) in want.items():
        assert O.fm_deemphasis(tau) == (b0, a1)
        assert fd.fm_deemphasis(tau) == (b0, a1)


def test_iir_oracle_vs_lfilter():
    u = np.random.default_rng(1)
    b = (u.standard_normal(3) + 1j * u.standard_normal(3)).astype(np.complex64)
    a = np.array([1.0, -0.5 + 0.2j, 0.1], dtype=np.complex64)
    x = cwgn(2, 5000)
    want = sps.lfilter(b.astype(np.complex128), a.astype(np.complex128), x.astype(np.complex128))
    got = np.concatenate(chunked(O.Iir(b, a).apply, x, [1, 2, 300, 4697]))
    assert np.max(np.abs(got - want)) <= 1e-5 * np.sqrt(np.mean(np.abs(want) ** 2))
    # a[0] != 1 divides through (scipy does the same)
    a2 = (2 * a).astype(np.complex64)
    got2 = O.Iir(b, a2).apply(x)
    want2 = sps.lfilter(b.astype(np.complex128), a2.astype(np.complex128), x.astype(np.complex128))
    assert np.max(np.abs(got2 - want2)) <= 1e-5 * np.sqrt(np.mean(np.abs(want2) ** 2))
    # a planted mistake (the sign of the feedback) is caught
    bad = sps.lfilter(b.astype(np.complex128), np.array([1.0, 0.5 - 0.2j, -0.1]), x.astype(np.complex128))
    assert np.max(np.abs(got - bad)) > 1e-2


@needs_ref
def test_iir_oracle_vs_reference_live():
    b0, a1 = O.fm_deemphasis(75e-6)
    x = np.random.default_rng(0).standard_normal(60_000).astype(np.float32)
    sizes = [17, 29_983, 1, 2, 30_000 - 3]
    a = np.concatenate(chunked(O.Iir([b0], [1, a1]).apply_f32, x, sizes))
    b = np.concatenate(chunked(O.RefIir([b0], [1, a1]).apply_f32, x, sizes))
    np.testing.assert_array_equal(a, b)
    bb = np.array([0.2 + 0.1j, 0.3 - 0.2j, 0.1j], np.complex64)
    aa = np.array([2.0, 0.5 - 0.3j, 0.1], np.complex64)
    xc = cwgn(4, 20_000)
    np.testing.assert_array_equal(np.concatenate(chunked(O.Iir(bb, aa).apply, xc, [2, 9_998, 10_000])),
                                  np.concatenate(chunked(O.RefIir(bb, aa).apply, xc, [2, 9_998, 10_000])))


@needs_ref
def test_fixture_taps_are_the_reference_table():
    iso, mono = fm_taps()
    for got, (fin, fout) in ((iso, (1_140_000, 228_000)), (mono, (228_000, 45_600))):
        up, down, taps = O.ref_resample_fir_params(fin, fout)
        assert (got[0], got[1]) == (up, down)
        np.testing.assert_array_equal(got[2], taps)


@needs_ref
@pytest.mark.parametrize("europe", [False, True])
def test_fm_oracle_vs_reference_demod_mono(europe):
    iso, mono = fm_taps()
    offset = 150_000
    n = FS // 2
    x = fm_input(5, n, -offset)
    sizes = [65_536, 131_072, 1, 50_000, n - 246_609]
    tau = 50e-6 if europe else 75e-6
    mine = chunked(lambda c: O.Fm(FS, offset, tau, iso, mono).demod(c)[0], x, sizes[:1])  # fresh object
    f, r = O.Fm(FS, offset, tau, iso, mono), O.RefFm(FS, offset, europe)
    got = chunked(lambda c: f.demod(c)[0], x, sizes)
    want = chunked(r.demod, x, sizes)
    for k, (g, w) in enumerate(zip(got, want)):
        w = np.zeros(0, np.int16) if w is None else w
        np.testing.assert_array_equal(g, w, err_msg=f"chunk {k}")
    assert sum(g.size for g in got) == 22_800  # 570,000 inputs, decimated by 5 twice
    np.testing.assert_array_equal(mine[0], got[0])
