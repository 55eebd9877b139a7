"""Deterministic synthetic signals and filter taps (numpy only).

* ``splitmix64`` -> uniform (0,1) -> Box-Muller complex white Gaussian noise,
  so any host (or a C port) reproduces the same samples bit for bit from a
  seed and a sample offset.
* ``kaiser_sinc``: windowed-sinc lowpass, the same design as
  scipy.signal.firwin(M, 2*cutoff, window=('kaiser', beta)) (unit DC gain),
  times ``gain``.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Outputs offset .. offset+n-1 of the splitmix64 stream seeded with ``seed``."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Uniform doubles in the open interval (0, 1)."""
    z = splitmix64(seed, n, offset) >> np.uint64(11)
    return (z.astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def cwgn(seed: int, n: int, offset: int = 0, chunk: int = 1 << 22) -> np.ndarray:
    """Unit-variance complex white Gaussian noise (each of I, Q has variance 1/2),
    complex64. Sample i uses uniforms 2i and 2i+1 of the stream."""
    out = np.empty(n, dtype=np.complex64)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        u = uniform(seed, 2 * m, 2 * (offset + s))
        r = np.sqrt(-np.log(u[0::2]))  # sqrt(-2 ln u) / sqrt(2)
        th = 2.0 * np.pi * u[1::2]
        out[s:s + m].real = r * np.cos(th)
        out[s:s + m].imag = r * np.sin(th)
    return out


def tone(n: int, freq_cycles_per_sample: float, amplitude: float = 1.0, phase: float = 0.0) -> np.ndarray:
    """amplitude * exp(j(2 pi f n + phase)) computed in FP64, returned as complex64.
    The phase index is reduced exactly for rational-looking frequencies by
    working in revolutions."""
    k = np.arange(n, dtype=np.float64)
    rev = np.mod(k * freq_cycles_per_sample, 1.0)
    return (amplitude * np.exp(1j * (2 * np.pi * rev + phase))).astype(np.complex64)


def tones_sum(n: int, freqs_cycles_per_sample, amplitudes, block: int = 1 << 16) -> np.ndarray:
    """sum_c a_c exp(j 2 pi f_c k) for k < n, complex128, as a product of
    per-block phasors (exact mod-1 phase reduction) with per-offset phasors."""
    f = np.asarray(freqs_cycles_per_sample, dtype=np.float64)
    a = np.asarray(amplitudes, dtype=np.complex128)
    nb = -(-n // block)
    m = np.arange(block, dtype=np.float64)
    base = np.exp(2j * np.pi * np.mod(np.outer(f, m), 1.0))                  # (C, block)
    starts = np.arange(nb, dtype=np.float64) * block
    blk = a[None, :] * np.exp(2j * np.pi * np.mod(np.outer(starts, f), 1.0))  # (nb, C)
    return (blk @ base).reshape(-1)[:n]


def kaiser_sinc(num_taps: int, cutoff: float, beta: float, gain: float = 1.0) -> np.ndarray:
    """Windowed-sinc lowpass: ``cutoff`` in cycles/sample (0.5 = Nyquist),
    Kaiser window of shape ``beta``, normalised to unit DC gain, times gain."""
    m = np.arange(num_taps, dtype=np.float64) - (num_taps - 1) / 2.0
    h = 2.0 * cutoff * np.sinc(2.0 * cutoff * m) * np.kaiser(num_taps, beta)
    h = h / h.sum() * gain
    return h.astype(np.float32)


def fm_signal(n: int, fs: float, tones_hz, deviation_hz: float, amplitude: float = 1.0,
              carrier_hz: float = 0.0) -> np.ndarray:
    """Complex baseband FM: message = mean of unit cosines at ``tones_hz``
    (so |message| <= 1), instantaneous frequency carrier + deviation*message.
    The phase is the exact running sum in FP64."""
    t = np.arange(n, dtype=np.float64)
    msg = np.zeros(n, dtype=np.float64)
    for f in tones_hz:
        msg += np.cos(2 * np.pi * (f / fs) * t)
    msg /= max(len(tones_hz), 1)
    inst = (carrier_hz + deviation_hz * msg) / fs  # cycles per sample
    ph = np.cumsum(inst) - inst[0]
    ph = np.mod(ph, 1.0)
    return (amplitude * np.exp(2j * np.pi * ph)).astype(np.complex64), msg
