"""The benchmark workloads, with their sizes, as concrete seeded inputs.

resampler_1ch   plain single-channel polyphase resampler: 2^20 samples of
                unit CWGN (seed 0) + a unit tone at 0.1 fs; up 4, down 5;
                96 Kaiser-sinc taps (beta 8, cutoff 0.5/5 cycles/sample at the
                up-rate, gain = up); 16 chunks of 2^16.
nco_16shift     2^24 samples CWGN (seed 1) at 25 Msps; 16 shifts
                round(-10e6 + k*20e6/15) Hz; 64 chunks of 2^18.
chan_32         2^24 samples at 20 Msps: CWGN (seed 2) + 32 unit tones, one
                inside each channel; 32 channels with centres U(+-9 MHz)
                (rounded to Hz, nonzero) and output rates drawn from
                {12.5k, 25k, 100k, 250k, 1M, 2.5M}; up/down by gcd; Kaiser taps
                of 32*up (beta 8, cutoff 0.5/max(up, down), gain up).
fir_bank_256    256 channels x 2^20 CWGN (seed 3); 512 real Kaiser-sinc taps
                per channel, cutoff U[0.05, 0.45].
iir_order4      2^24 samples CWGN (seed 6), complex float, through a 4th-order
                Butterworth low-pass at 0.1 cycles/sample (bilinear
                transform, b and a as scipy.signal.butter gives them); 16
                chunks of 2^20.
serving         Poisson arrivals (2000/s) of single-channel streams at
                25 Msps drawn as in chan_32 with the shift on nco_16shift's
                +-10 MHz grid; chunks U[4k, 256k]; a share of the streams
                carries 3-tone FM (300-3000 Hz, 75 kHz deviation) and ends in
                the polar discriminator at 250 ksps.

Parameter draws use their own splitmix64 streams (seed + 1000 + purpose) so
changing one draw never shifts the samples.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

from .signals import cwgn, fm_signal, kaiser_sinc, tone, tones_sum, uniform

RATES = [12_500, 25_000, 100_000, 250_000, 1_000_000, 2_500_000]


def _gcd(a, b):
    return math.gcd(int(a), int(b))


def channel(fs: int, shift_hz: int, out_rate: int, taps_per_up: int = 32, beta: float = 8.0, tail=None) -> dict:
    g = _gcd(fs, out_rate)
    up, down = out_rate // g, fs // g
    taps = kaiser_sinc(taps_per_up * up, 0.5 / max(up, down), beta, gain=up).astype(np.complex64)
    d = dict(output_sample_rate_in_sps=int(out_rate), freq_shift_in_hz=int(shift_hz), up_rate=int(up),
             down_rate=int(down), taps=taps)
    if tail:
        d["tail"] = tail
    return d


# --------------------------------------------------------------------------
@dataclass
class Resampler1ch:
    n: int = 1 << 20
    chunk: int = 1 << 16
    up: int = 4
    down: int = 5

    def input(self) -> np.ndarray:
        x = cwgn(0, self.n).astype(np.complex128) + tone(self.n, 0.1).astype(np.complex128)
        return x.astype(np.complex64)

    def taps(self) -> np.ndarray:
        return kaiser_sinc(24 * self.up, 0.5 / max(self.up, self.down), 8.0, gain=self.up).astype(np.complex64)


@dataclass
class Nco16:
    n: int = 1 << 24
    chunk: int = 1 << 18
    fs: int = 25_000_000

    def shifts(self) -> List[int]:
        return [int(round(-10e6 + k * 20e6 / 15)) for k in range(16)]

    def input(self) -> np.ndarray:
        return cwgn(1, self.n)

    def input_ci16(self, scale: float = 4096.0) -> np.ndarray:
        """The same noise as complex int16 (an (n, 2) int16 array), rounded from
        `scale` times the float samples: an SDR front end's sample format."""
        x = self.input()
        v = np.stack([np.round(x.real * scale), np.round(x.imag * scale)], axis=1)
        return v.clip(-32768, 32767).astype(np.int16)


@dataclass
class IirOrder4:
    n: int = 1 << 24
    chunk: int = 1 << 20

    def coeffs(self):
        from scipy.signal import butter
        b, a = butter(4, 0.2)  # cutoff 0.2 of Nyquist = 0.1 cycles/sample
        return b.astype(np.complex64), a.astype(np.complex64)

    def input(self) -> np.ndarray:
        return cwgn(6, self.n)


@dataclass
class FmDemod:
    """FM mono demodulation of an RTL-SDR style stream: 1.14 Msps, the station
    200 kHz off DC, 1,140,000 -> 228,000 -> 45,600 sps through 64-tap
    filters (Kaiser-windowed here; the reference uses its REMEZ table)."""
    n: int = 1 << 24
    chunk: int = 1 << 18
    fs: int = 1_140_000
    offset: int = -200_000

    def taps(self):
        t = kaiser_sinc(64, 0.5 / 5, 8.0).astype(np.complex64)
        return (1, 5, t), (1, 5, t)

    def input(self) -> np.ndarray:
        from .signals import fm_signal
        sig, _ = fm_signal(self.n, self.fs, [400.0, 1300.0, 2700.0], 75e3, carrier_hz=-self.offset)
        return (sig + 0.02 * cwgn(7, self.n)).astype(np.complex64)


@dataclass
class Chan32:
    n: int = 1 << 24
    fs: int = 20_000_000
    n_chan: int = 32
    seed: int = 2

    def channels(self) -> List[dict]:
        u = uniform(self.seed + 1000, 3 * self.n_chan)
        chans = []
        for c in range(self.n_chan):
            f = int(round(-9e6 + 18e6 * u[3 * c]))
            if f == 0:
                f = 1
            rate = RATES[min(int(u[3 * c + 1] * len(RATES)), len(RATES) - 1)]
            chans.append(channel(self.fs, f, rate))
        return chans

    def input(self) -> np.ndarray:
        u = uniform(self.seed + 1000, 3 * self.n_chan)
        freqs = []
        for c, ch in enumerate(self.channels()):
            # one tone inside each channel's passband: the shift moves -f_c to DC
            off = (u[3 * c + 2] - 0.5) * 0.5 * ch["output_sample_rate_in_sps"]
            freqs.append((-ch["freq_shift_in_hz"] + off) / self.fs)
        x = tones_sum(self.n, freqs, np.ones(len(freqs)))
        x += cwgn(self.seed, self.n)
        return x.astype(np.complex64)

    def total_outputs(self) -> int:
        return sum(-(-self.n * c["up_rate"] // c["down_rate"]) for c in self.channels())


@dataclass
class FirBank256:
    n: int = 1 << 20
    n_chan: int = 256
    n_taps: int = 512

    def taps(self) -> np.ndarray:
        u = uniform(3 + 1000, self.n_chan)
        return np.stack([kaiser_sinc(self.n_taps, 0.05 + 0.4 * u[c], 8.0) for c in range(self.n_chan)])

    def input(self, chan: int) -> np.ndarray:
        return cwgn(3, self.n, offset=chan * self.n)


@dataclass
class ServingRequest:
    t_arrival: float
    stream: int
    chunk: int


@dataclass
class Serving:
    fs: int = 25_000_000
    rate_hz: float = 2000.0
    max_streams: int = 200
    n_requests: int = 2000
    chunk_lo: int = 4096
    chunk_hi: int = 256 * 1024
    fm_share: float = 0.25
    hbm_budget: int = 64 << 30
    seed: int = 5
    streams: List[dict] = field(default_factory=list)

    def stream_channel(self, s: int) -> dict:
        u = uniform(self.seed + 2000, 4, offset=4 * s)
        if u[0] < self.fm_share:
            # FM stream: 250 ksps output, discriminator tail
            f = int(round(-10e6 + 20e6 * u[1]))
            return channel(self.fs, f or 1, 250_000, tail="discriminator")
        f = int(round(-10e6 + round(u[1] * 15) * 20e6 / 15))
        rate = RATES[min(int(u[2] * len(RATES)), len(RATES) - 1)]
        return channel(self.fs, f or 1, rate)

    def requests(self) -> List[ServingRequest]:
        u = uniform(self.seed + 3000, 3 * self.n_requests)
        t = 0.0
        reqs = []
        for i in range(self.n_requests):
            t += -math.log(u[3 * i]) / self.rate_hz
            s = int(u[3 * i + 1] * self.max_streams)
            m = self.chunk_lo + int(u[3 * i + 2] * (self.chunk_hi - self.chunk_lo + 1))
            reqs.append(ServingRequest(t, s, m))
        return reqs

    def stream_input(self, s: int, offset: int, n: int) -> np.ndarray:
        """Samples offset..offset+n of stream s: CWGN at -20 dB plus, for FM
        streams, the FM carrier at the channel's centre."""
        ch = self.stream_channel(s)
        x = 0.1 * cwgn(self.seed * 7919 + s, n, offset=offset).astype(np.complex128)
        if ch.get("tail"):
            sig, _ = fm_signal(offset + n, self.fs, self.fm_tones(s), 75e3, carrier_hz=-ch["freq_shift_in_hz"])
            x += sig[offset:].astype(np.complex128)
        return x.astype(np.complex64)

    def fm_tones(self, s: int):
        u = uniform(self.seed + 4000, 3, offset=3 * s)
        return [300.0 + 2700.0 * v for v in u]
