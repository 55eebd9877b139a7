"""Synthetic inputs and workload definitions (libfalcon_dsp_amd.signals,
libfalcon_dsp_amd.workloads): deterministic, and the tap designs agree with
scipy's window-method FIR design."""
import numpy as np
from scipy import signal

from libfalcon_dsp_amd.signals import cwgn, fm_signal, kaiser_sinc, tone, tones_sum, uniform
from libfalcon_dsp_amd.workloads import Chan32, FirBank256, Nco16, Resampler1ch, channel


def test_streams_are_deterministic():
    assert np.array_equal(cwgn(5, 1000), cwgn(5, 1000))
    assert not np.array_equal(cwgn(5, 1000), cwgn(6, 1000))
    assert np.array_equal(uniform(9, 10), uniform(9, 10))
    u = uniform(1, 100_000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01


def test_cwgn_is_unit_variance_complex():
    x = cwgn(0, 1 << 18).astype(np.complex128)
    assert abs(np.mean(np.abs(x) ** 2) - 1.0) < 0.01
    assert abs(np.mean(x)) < 0.01
    assert abs(np.var(x.real) - 0.5) < 0.01 and abs(np.var(x.imag) - 0.5) < 0.01


def test_kaiser_sinc_matches_firwin():
    for n, cutoff, beta, gain in [(96, 0.1, 8.0, 4.0), (32, 0.5 / 80, 8.0, 1.0), (33, 0.25, 5.0, 2.0)]:
        h = kaiser_sinc(n, cutoff, beta, gain)
        # firwin's cutoff is relative to Nyquist (fs = 2 -> 2 * cycles/sample)
        want = signal.firwin(n, 2 * cutoff, window=("kaiser", beta)) * gain
        np.testing.assert_allclose(h, want, rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(h, h[::-1], atol=1e-7)


def test_tones_sum_matches_direct_sum():
    f = [0.01, -0.2, 0.3333]
    a = [1.0, 0.5, 2.0]
    x = tones_sum(5000, f, a)
    want = sum(ai * np.exp(2j * np.pi * fi * np.arange(5000)) for fi, ai in zip(f, a))
    assert np.max(np.abs(x - want)) < 1e-4
    assert np.max(np.abs(tone(100, 0.25) - np.exp(0.5j * np.pi * np.arange(100)))) < 1e-6


def test_fm_signal_phase_follows_message():
    x, msg = fm_signal(2000, 250_000.0, [1000.0], 50_000.0)
    d = np.angle(x[1:] * np.conj(x[:-1]))
    assert np.max(np.abs(d - 2 * np.pi * 50_000.0 * msg[1:] / 250_000.0)) < 1e-4


def test_channel_descriptor_rates():
    d = channel(20_000_000, 1_000_000, 250_000)
    assert (d["up_rate"], d["down_rate"]) == (1, 80) and d["taps"].size == 32
    d = channel(1_000_000, -5, 800_000)
    assert (d["up_rate"], d["down_rate"]) == (4, 5) and d["taps"].size == 128


def test_workload_definitions():
    w = Chan32()
    chans = w.channels()
    assert len(chans) == 32 and w.n == 1 << 24 and w.fs == 20_000_000
    downs = sorted(c["down_rate"] for c in chans)
    assert set(downs) <= {8, 20, 80, 200, 800, 1600} and all(c["up_rate"] == 1 for c in chans)
    assert all(abs(c["freq_shift_in_hz"]) <= 9_000_000 and c["freq_shift_in_hz"] != 0 for c in chans)
    assert w.total_outputs() == sum(-(-w.n // c["down_rate"]) for c in chans)
    r = Resampler1ch()
    assert (r.up, r.down, r.taps().size, r.n // r.chunk) == (4, 5, 96, 16)
    assert len(Nco16().shifts()) == 16
    fb = FirBank256(n=1024)
    assert fb.taps().shape == (256, 512)
