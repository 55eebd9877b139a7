"""The oracle (oracle/fdsp_oracle.c, a C restatement of the reference's CPU
algorithms; test infrastructure only) pinned on the CPU against:

- the known-answer vectors the reference's own unit tests hold
  (tests/golden/kat_reference.json, source file:line per entry);
- outputs of the reference's compiled CPU classes on seeded inputs streamed
  in uneven chunks (tests/golden/ref_vectors.npz, written by
  tests/golden/make_ref_vectors.py from oracle/_ref), and oracle/_ref live
  where it is built;
- scipy (upfirdn, lfilter), FP64 closed forms of the NCO and the
  discriminator, and invariants (|y| = |x|, chunk invariance, composition).

Tolerances: counts, rollover indices and resampler state exact; samples
within 1e-6 relative to the input RMS against the compiled reference (the
same float32 arithmetic, possibly contracted differently by the compiler),
the stated float32 bounds against FP64 forms.
"""
import json
import os

import numpy as np
import pytest
from scipy import signal

from libfalcon_dsp_amd.signals import cwgn, kaiser_sinc
from oracle import oracle as O
from tests.helpers import chunks, max_err, rms

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KAT = json.load(open(os.path.join(GOLDEN, "kat_reference.json")))
REF = np.load(os.path.join(GOLDEN, "ref_vectors.npz"))
CHUNKS = [int(c) for c in REF["chunks"]]
N = int(REF["n"][0])
REF_TOL = 1e-6


def _stream(obj, x, sizes=None):
    return [obj.apply(x[lo:hi]) for lo, hi in chunks(x.size, sizes or CHUNKS)]


# --------------------------------------------------------------------------
# reference KATs
# --------------------------------------------------------------------------
def test_gcd_lcm_rat_approx_kats():
    for a, b, g in KAT["gcd"]["cases"]:
        assert O.gcd(a, b) == g
    for a, b, m in KAT["lcm"]["cases"]:
        assert O.lcm(a, b) == m
    for f, md, n, d in KAT["rat_approx"]["cases"]:
        assert O.rat_approx(f, md) == (n, d), (f, md)


@pytest.mark.parametrize("case", KAT["resampler"]["cases"], ids=lambda c: c["name"])
def test_resampler_kats(case):
    h = np.asarray(case["h"], np.complex64)
    x = np.ones(case["x_ones"], np.complex64)
    r = O.Resampler(case["up"], case["down"], h)
    y = r.apply(x)
    assert y.size == -(-x.size * case["up"] // case["down"])
    want = list(case["expected_head"])
    if "expected_fill" in case:
        want += [case["expected_fill"]] * (case["expected_len"] - len(want))
    n = min(y.size, len(want))
    # the reference tests compare at 0.1; these inputs are exact in float32
    np.testing.assert_allclose(y[:n].real, want[:n], atol=1e-5)
    np.testing.assert_allclose(y[:n].imag, 0.0, atol=1e-6)


@pytest.mark.parametrize("case", KAT["fir"]["cases"], ids=lambda c: c["name"])
def test_fir_kats(case):
    f = O.Fir(np.asarray(KAT["fir"]["h"], np.complex64))
    x = np.asarray(case["x"], np.complex64)
    if case.get("one_at_a_time"):
        y = np.concatenate([f.apply(x[i:i + 1]) for i in range(x.size)])
    else:
        y = f.apply(x)
    np.testing.assert_array_equal(y.real, case["expected"])
    np.testing.assert_array_equal(y.imag, 0.0)


def test_discriminator_kat():
    k = KAT["discriminator"]
    x = np.asarray([complex(a, b) for a, b in k["x"]], np.complex64)
    np.testing.assert_allclose(O.Discriminator().apply(x), k["expected"], atol=k["tol"])


# --------------------------------------------------------------------------
# compiled-reference vectors (tests/golden/ref_vectors.npz)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("i", range(4))
def test_freq_shift_vs_reference_vectors(i):
    fs, f, seed = (int(v) for v in REF[f"fs{i}_params"])
    x = cwgn(seed, N)
    y = np.concatenate(_stream(O.FreqShift(fs, f), x))
    want = REF[f"fs{i}_y"]
    assert y.size == want.size
    assert max_err(y, want) <= REF_TOL * rms(x)


@pytest.mark.parametrize("i", range(6))
def test_resampler_vs_reference_vectors(i):
    up, down, seed = (int(v) for v in REF[f"rs{i}_params"])
    x = cwgn(seed, N)
    ys = _stream(O.Resampler(up, down, REF[f"rs{i}_taps"]), x)
    assert [y.size for y in ys] == list(REF[f"rs{i}_counts"])
    assert max_err(np.concatenate(ys), REF[f"rs{i}_y"]) <= REF_TOL * rms(x)


@pytest.mark.parametrize("i", range(3))
def test_fir_vs_reference_vectors(i):
    x = cwgn(int(REF[f"fir{i}_seed"][0]), N)
    y = np.concatenate(_stream(O.Fir(REF[f"fir{i}_taps"]), x))
    assert max_err(y, REF[f"fir{i}_y"]) <= REF_TOL * rms(x)


def test_discriminator_vs_reference_vectors():
    x = cwgn(int(REF["disc_seed"][0]), N)
    y = np.concatenate(_stream(O.Discriminator(), x))
    want = REF["disc_y"]
    assert y.size == want.size == N - 1
    d = np.angle(np.exp(1j * (y.astype(np.float64) - want)))
    assert np.max(np.abs(d)) <= 1e-5


def test_channelizer_vs_reference_vectors():
    fs, seed = (int(v) for v in REF["chan_fs"])
    descs = []
    for c in range(5):
        rate, shift, up, down = (int(v) for v in REF[f"chan{c}_params"])
        descs.append(dict(output_sample_rate_in_sps=rate, freq_shift_in_hz=shift, up_rate=up, down_rate=down,
                          taps=REF[f"chan{c}_taps"]))
    x = cwgn(seed, 4 * N)
    ch = O.Channelizer(fs, descs)
    per = [[] for _ in descs]
    for lo, hi in chunks(x.size, CHUNKS):
        for c, y in enumerate(ch.apply(x[lo:hi])):
            per[c].append(y)
    for c in range(5):
        assert [y.size for y in per[c]] == list(REF[f"chan{c}_counts"]), f"chan {c} counts"
        assert max_err(np.concatenate(per[c]), REF[f"chan{c}_y"]) <= REF_TOL * rms(x), f"chan {c}"


# --------------------------------------------------------------------------
# scipy and closed forms
# --------------------------------------------------------------------------
@pytest.mark.parametrize("up,down,m", [(1, 2, 3), (4, 5, 96), (3, 7, 50), (2, 1, 20), (1, 8, 33), (7, 3, 40),
                                       (1, 1, 9)])
def test_resampler_vs_upfirdn(up, down, m):
    # one-shot output is the first ceil(N*up/down) samples of upfirdn
    h = kaiser_sinc(m, 0.5 / max(up, down), 6.0, gain=up).astype(np.float64)
    h[0] += 0.3  # break the symmetry: a reversed filter must not pass
    x = cwgn(7, 3000)
    y = O.Resampler(up, down, h.astype(np.complex64)).apply(x)
    want = signal.upfirdn(h, x.astype(np.complex128), up, down)[:y.size]
    assert y.size == -(-x.size * up // down)
    assert max_err(y, want) <= 2e-6 * rms(x) * np.abs(h).sum()
    # the check discriminates: reversed taps or a one-sample shift fail it
    bad = signal.upfirdn(h[::-1], x.astype(np.complex128), up, down)[:y.size]
    assert max_err(y, bad) > 1e-2 * rms(x)
    assert max_err(y[1:], want[:-1]) > 1e-2 * rms(x)


def test_fir_vs_lfilter():
    h = (cwgn(11, 37) / 6).astype(np.complex64)
    x = cwgn(12, 5000)
    y = np.concatenate(_stream(O.Fir(h), x, [1, 999, 3, 4000]))
    want = signal.lfilter(h.astype(np.complex128), [1.0], x.astype(np.complex128))
    assert max_err(y, want) <= 2e-6 * rms(x) * np.abs(h).sum()


@pytest.mark.parametrize("fs,f", [(1_000_000, 23_000), (25_000_000, -9_876_543), (20_000_000, 1), (48_000, 12_000),
                                  (100_000, 1), (100_000, -7)])
def test_freq_shift_closed_form(fs, f):
    r, w = O.freq_shift_params(fs, f)
    # R: fs/|f| scaled by ten until it reaches 1e5, rounded (freq_shift.cc:107-122)
    ratio = fs / abs(f)
    assert r >= 1e5 and (ratio >= 1e5 or r / 10 < 1e5 + 1)
    assert abs(w - 2 * np.pi * f / fs) <= 1e-15 * abs(w) + 1e-18
    x = cwgn(5, 3 * r // 2 if r < 200_000 else 4096)
    y = np.concatenate(_stream(O.FreqShift(fs, f), x, [1000, 1, 70_001]))
    want = O.fp64_freq_shift(x.astype(np.complex128), fs, f, rollover=r)
    assert max_err(y, want) <= 2e-6 * rms(x)
    np.testing.assert_allclose(np.abs(y), np.abs(x), rtol=2e-6, atol=1e-7)


def test_freq_shift_rollover_index_is_exact():
    fs, f = 1_000_000, 23_000
    r, _ = O.freq_shift_params(fs, f)
    s = O.FreqShift(fs, f)
    done = 0
    for m in (r - 5, 3, 7, 2 * r + 11):
        s.apply(cwgn(1, m))
        done += m
        assert s.samples_handled == done % r


def test_discriminator_closed_form_fm():
    # 2 pi * deviation * message / fs with no phase wrap (deviation/fs = 0.3)
    from libfalcon_dsp_amd.signals import fm_signal
    fs = 250_000
    x, msg = fm_signal(20_000, fs, [300.0, 1100.0, 2900.0], 75e3)
    y = np.concatenate(_stream(O.Discriminator(), x, [1, 5000, 14999]))
    want = 2 * np.pi * 75e3 * msg[1:] / fs
    assert np.max(np.abs(y - want)) <= 2e-4


def test_channelizer_is_shift_then_resample():
    fs = 2_000_000
    descs = []
    for rate, shift in ((250_000, 60_000), (500_000, -200_000), (300_000, 123_457)):
        g = int(np.gcd(fs, rate))
        up, down = rate // g, fs // g
        taps = kaiser_sinc(21 * up, 0.5 / max(up, down), 8.0, gain=up).astype(np.complex64)
        descs.append(dict(output_sample_rate_in_sps=rate, freq_shift_in_hz=shift, up_rate=up, down_rate=down,
                          taps=taps))
    x = cwgn(8, 20_000)
    got = O.Channelizer(fs, descs).apply(x)
    for d, y in zip(descs, got):
        r, _ = O.freq_shift_params(fs, d["freq_shift_in_hz"])
        sh = O.fp64_freq_shift(x.astype(np.complex128), fs, d["freq_shift_in_hz"], rollover=r)
        want = O.fp64_resample(sh, d["taps"].astype(np.complex128), d["up_rate"], d["down_rate"], n_out=y.size)
        assert max_err(y, want) <= 1e-5 * rms(x)


@pytest.mark.parametrize("up,down", [(4, 5), (1, 8), (3, 7), (5, 2)])
def test_resampler_chunk_invariance(up, down):
    h = kaiser_sinc(12 * max(up, down), 0.5 / max(up, down), 7.0, gain=up).astype(np.complex64)
    x = cwgn(9, 7001)
    one = O.Resampler(up, down, h)
    y1 = one.apply(x)
    many = O.Resampler(up, down, h)
    y2 = np.concatenate(_stream(many, x, [1, 2, 3, 5, 1000, 7, 13]))
    assert y1.size == y2.size
    assert max_err(y1, y2) <= 1e-6 * rms(x)
    assert one.state_fields == many.state_fields


# --------------------------------------------------------------------------
# live compiled reference, where it is built (this container)
# --------------------------------------------------------------------------
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")


@needs_ref
def test_rate_helpers_vs_reference_live():
    import ctypes as C
    L = O.ref()
    u = np.random.default_rng(3)
    for _ in range(200):
        a, b = (int(v) for v in u.integers(1, 10**9, 2))
        assert O.gcd(a, b) == L.ref_gcd(a, b)
        assert O.lcm(a % 100000 + 1, b % 100000 + 1) == L.ref_lcm(a % 100000 + 1, b % 100000 + 1)
        f, md = float(u.uniform(0, 3)), int(u.integers(1, 2000))
        n, d = C.c_int64(), C.c_int64()
        L.ref_rat_approx(f, md, C.byref(n), C.byref(d))
        assert O.rat_approx(f, md) == (n.value, d.value)
        fs, fh = int(u.integers(1000, 2**31)), int(u.integers(-2**30, 2**30))
        r, w = C.c_uint32(), C.c_double()
        L.ref_freq_shift_params(fs, fh, C.byref(r), C.byref(w))
        assert O.freq_shift_params(fs, fh) == (r.value, w.value)


@needs_ref
@pytest.mark.parametrize("seed", range(4))
def test_resampler_vs_reference_live(seed):
    u = np.random.default_rng(seed)
    up, down = int(u.integers(1, 9)), int(u.integers(1, 12))
    h = kaiser_sinc(int(u.integers(1, 80)), 0.5 / max(up, down), 6.0, gain=up).astype(np.complex64)
    x = cwgn(seed + 50, 9000)
    sizes = [int(v) for v in u.integers(1, 3000, 6)]
    a = _stream(O.Resampler(up, down, h), x, sizes)
    b = _stream(O.RefResampler(up, down, h), x, sizes)
    assert [y.size for y in a] == [y.size for y in b]
    assert max_err(np.concatenate(a), np.concatenate(b)) <= REF_TOL * rms(x)
