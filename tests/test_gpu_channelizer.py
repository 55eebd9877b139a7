"""Multi-rate channelizer on the GPU, both forms, vs the oracle (reference
src/transform/falcon_dsp_multi_rate_channelizer.cc:81-173, _cuda.cu:375-577).

Bound: max |err| <= 1e-4 * RMS(x) per channel against the oracle's
shift-then-resample; output counts, NCO index and (coeff_phase, xOffset)
exactly equal after every chunk.
"""
import numpy as np
import pytest

import libfalcon_dsp_amd as fd
from libfalcon_dsp_amd.signals import cwgn
from libfalcon_dsp_amd.workloads import Chan32, channel
from oracle import oracle as O
from tests.helpers import assert_close_rms, chunks, rms

pytestmark = pytest.mark.gpu

# agreement between work splits of the batched kernel, relative to input RMS
TILING_TOL = 2e-6


def _chans_small(fs=20_000_000):
    return [
        channel(fs, 1_234_567, 250_000),
        channel(fs, -3_000_001, 2_500_000),
        channel(fs, 7_777_777, 12_500),
        channel(fs, -8_999_999, 1_000_000),
        channel(fs, 23, 100_000),
        channel(fs, 4_400_000, 25_000),
        channel(fs, -1, 3_000_000),      # up 3, down 20
        channel(fs, 2_000_000, 1_600_000),  # up 2, down 25
    ]


def _compare(fs, chans, x, sizes, batched):
    gpu = fd.MultiRateChannelizer(fs, chans, batched=batched)
    ref = O.Channelizer(fs, chans)
    xr = rms(x)
    for lo, hi in chunks(x.size, sizes):
        ys = gpu.apply(x[lo:hi])
        rs = ref.apply(x[lo:hi])
        for c, (y, r) in enumerate(zip(ys, rs)):
            assert y.size == r.size, f"chan {c} chunk [{lo},{hi}): {y.size} vs {r.size}"
            assert_close_rms(y, r, xr, what=f"chan {c} chunk [{lo},{hi}) batched={batched}")
            k, ph, xo, it, ot = gpu.state(c)
            rk, rph, rxo = ref.state(c)
            assert (ph, xo) == (rph, rxo), f"chan {c}: resampler state after [{lo},{hi})"
            if chans[c]["freq_shift_in_hz"]:
                assert k == rk, f"chan {c}: NCO index after [{lo},{hi})"
            assert it == hi
    return gpu


@pytest.mark.parametrize("batched", [True, False])
def test_small_channel_set_chunked(batched):
    fs = 20_000_000
    x = cwgn(2, 400_000)
    _compare(fs, _chans_small(fs), x, [65_536, 1, 3, 100_003, 17], batched)


@pytest.mark.parametrize("batched", [True, False])
def test_rollover_crossing_windows(batched):
    # R = 1e5..1e6: a 1.2M-sample stream crosses several rollovers, and the
    # chunk boundaries below put windows astride them
    fs = 1_000_000
    chans = [channel(fs, 23_000, 50_000), channel(fs, -7, 200_000), channel(fs, 333_333, 250_000)]
    x = cwgn(17, 1_200_000)
    _compare(fs, chans, x, [434_780, 5, 1, 262_144], batched)


@pytest.mark.parametrize("batched", [True, False])
def test_empty_chunk_keeps_the_stream(batched):
    # the reference's apply pushes one (here empty) output per channel and an
    # empty chunk moves no state (falcon_dsp_multi_rate_channelizer.cc:146-164)
    fs = 20_000_000
    chans = _chans_small(fs)
    x = cwgn(41, 50_000)
    gpu, ref = fd.MultiRateChannelizer(fs, chans, batched=batched), O.Channelizer(fs, chans)
    xr = rms(x)
    for lo, hi in ((0, 20_000), (20_000, 20_000), (20_000, 20_001), (20_001, 20_001), (20_001, 50_000)):
        print(f"channelizer batched={batched} chunk [{lo},{hi})", flush=True)
        ys, rs = gpu.apply(x[lo:hi]), ref.apply(x[lo:hi])
        assert len(ys) == len(rs) == len(chans)
        for c, (y, r) in enumerate(zip(ys, rs)):
            assert y.size == r.size, f"chan {c} chunk [{lo},{hi})"
            if r.size:
                assert_close_rms(y, r, xr, what=f"chan {c} chunk [{lo},{hi})")
            assert gpu.state(c)[3] == hi


def test_batched_equals_plain_and_state():
    fs = 20_000_000
    chans = _chans_small(fs)
    x = cwgn(3, 300_000)
    a = fd.MultiRateChannelizer(fs, chans, batched=True)
    b = fd.MultiRateChannelizer(fs, chans, batched=False)
    for lo, hi in chunks(x.size, [70_000, 11]):
        ya, yb = a.apply(x[lo:hi]), b.apply(x[lo:hi])
        for c in range(len(chans)):
            assert ya[c].size == yb[c].size
            assert_close_rms(ya[c], yb[c], rms(x))
            assert a.state(c) == b.state(c)


def _chans_grouped(fs=1_000_000):
    # channels sharing a rate form matrix-core groups: dense windows (down 20,
    # 25 and 4 < taps), sparse windows (down 80 >= 64 taps), windows padded
    # from 5 and 48 taps, an identity shift, and nine channels of one rate
    # (split into groups of 5 and 4)
    return ([channel(fs, f, 50_000) for f in (23_000, -7, 333_333, 0)]
            + [channel(fs, f, 12_500, taps_per_up=48) for f in (100_001, -250_000, 7)]
            + [channel(fs, f, 250_000, taps_per_up=5) for f in (1, -499_999, 125_000)]
            + [channel(fs, 3_000 * k + 11, 40_000) for k in range(9)])


def test_grouped_channels_chunked():
    # R = 1000 for the 23 kHz shift: windows straddle NCO rollovers often
    fs = 1_000_000
    x = cwgn(21, 700_000)
    _compare(fs, _chans_grouped(fs), x, [434_780, 5, 1, 131_071, 3], True)


# the channel-group kernel's work splits (FDSP_KERNEL_GROUPS_F32)
@pytest.mark.parametrize("seg,block", [(256, 128), (2048, 512), (16384, 256), (fd.TILING_STRIDED, 256),
                                       (fd.TILING_STRIDED, 512), (fd.TILING_CLAIM, 128), (fd.TILING_CLAIM, 512)])
def test_grouped_tilings_agree(seg, block):
    fs = 1_000_000
    chans = _chans_grouped(fs)
    x = cwgn(22, 200_000)
    b0 = fd.MultiRateChannelizer(fs, chans)
    b0.set_kernel("groups")
    base = b0.apply(x)
    g = fd.MultiRateChannelizer(fs, chans)
    g.set_kernel("groups")
    g.set_tiling(seg, block)
    # the work split only changes where the NCO phasor recurrence restarts
    # (a few float32 rounding steps), so tilings agree to ~1e-7
    for c, (a, b) in enumerate(zip(base, g.apply(x))):
        assert a.size == b.size
        assert_close_rms(b, a, rms(x), tol=TILING_TOL, what=f"chan {c}")


@pytest.mark.parametrize("chans_of", [_chans_small, _chans_grouped])
def test_alone_equals_in_the_mix(chans_of):
    # each channel's output does not depend on which other channels run (a
    # channel alone takes the one-channel path, in the mix it may share a
    # matrix-core group, so the two agree within the tolerance)
    fs = 20_000_000 if chans_of is _chans_small else 1_000_000
    chans = chans_of(fs)
    x = cwgn(4, 200_000)
    mix = fd.MultiRateChannelizer(fs, chans).apply(x)
    for c, ch in enumerate(chans):
        alone = fd.MultiRateChannelizer(fs, [ch]).apply(x)[0]
        assert alone.size == mix[c].size
        assert_close_rms(alone, mix[c], rms(x), what=f"chan {c}")


def test_reset_replays_identically():
    fs = 20_000_000
    chans = _chans_small(fs)
    x = cwgn(6, 150_000)
    g = fd.MultiRateChannelizer(fs, chans)
    first = g.apply(x)
    g.apply(x[:1000])
    g.reset_state()
    again = g.apply(x)
    for a, b in zip(first, again):
        np.testing.assert_array_equal(a, b)


def test_discriminator_tail():
    from libfalcon_dsp_amd.signals import fm_signal
    fs = 25_000_000
    ch = channel(fs, 2_000_000, 250_000, tail="discriminator")
    x, _ = fm_signal(500_000, fs, [400.0, 1100.0, 2900.0], 75e3, carrier_hz=-2_000_000)
    x = (x + 0.01 * cwgn(1, x.size)).astype(np.complex64)
    g = fd.MultiRateChannelizer(fs, [ch])
    ref = O.Channelizer(fs, [ch])
    dref = O.Discriminator()
    for lo, hi in chunks(x.size, [100_000, 3, 77_777]):
        got = g.apply(x[lo:hi])[0]
        want = dref.apply(ref.apply(x[lo:hi])[0])
        assert got.dtype == np.float32 and got.size == want.size
        d = np.angle(np.exp(1j * (got.astype(np.float64) - want)))
        assert d.size == 0 or np.max(np.abs(d)) <= 1e-3


@pytest.mark.parametrize("chans_of", [_chans_small, _chans_grouped])
def test_graph_replays_consecutive_chunks(chans_of):
    fs = 20_000_000 if chans_of is _chans_small else 1_000_000
    chans = chans_of(fs)
    chunk = 65_536
    x = cwgn(8, chunk * 5)
    g = fd.MultiRateChannelizer(fs, chans)
    d_in = fd.DeviceBuffer(chunk)
    outs = [fd.DeviceBuffer(m) for m in g.max_out(chunk)]
    g.graph_create(chunk, d_in, outs)
    ref = O.Channelizer(fs, chans)
    s = fd.Stream()
    for k in range(5):
        d_in.upload(x[k * chunk:(k + 1) * chunk], s)
        g.graph_launch(s)
        s.sync()
        counts = g.graph_counts()
        rs = ref.apply(x[k * chunk:(k + 1) * chunk])
        for c in range(len(chans)):
            assert counts[c] == rs[c].size
            assert_close_rms(outs[c].download(counts[c]), rs[c], rms(x), what=f"replay {k} chan {c}")


def test_workload_chan_32_slice():
    w = Chan32(n=1 << 20)
    chans = w.channels()
    x = w.input()
    g = fd.MultiRateChannelizer(w.fs, chans)
    ref = O.Channelizer(w.fs, chans)
    xr = rms(x)
    for lo, hi in chunks(x.size, [1 << 18]):
        ys, rs = g.apply(x[lo:hi]), ref.apply(x[lo:hi])
        for c in range(len(chans)):
            assert ys[c].size == rs[c].size
            assert_close_rms(ys[c], rs[c], xr, what=f"chan {c}")


def test_fused_groups_match_oracle():
    # opt-in bf16x2 form: two bf16-plane groups (down 8 and down 20) share
    # staged spans when fused; chunks of every size, including ones shorter
    # than a tile, and an unaligned input start (which takes the unfused form)
    w = Chan32(n=1 << 19)
    chans = [c for c in w.channels() if c["down_rate"] in (8, 20)]
    x = w.input()
    g = fd.MultiRateChannelizer(w.fs, chans, arith="bf16x2")
    on, members = g.get_fusion()
    assert not on and members == 2
    g.set_fusion(True)
    ref = O.Channelizer(w.fs, chans)
    xr = rms(x)
    for lo, hi in chunks(x.size, [100_003, 1_920, 7, 65_536, 1]):
        ys, rs = g.apply(x[lo:hi]), ref.apply(x[lo:hi])
        for c in range(len(chans)):
            assert ys[c].size == rs[c].size
            assert_close_rms(ys[c], rs[c], xr, what=f"chan {c} [{lo},{hi})")
    # device form on an odd sample offset (8-byte aligned input)
    n = 70_001
    d = fd.DeviceBuffer.from_array(x[: n + 1])
    g2 = fd.MultiRateChannelizer(w.fs, chans, arith="bf16x2")
    g2.set_fusion(True)
    outs = [fd.DeviceBuffer(m) for m in g2.max_out(n)]
    counts = g2.apply_device(d, n, outs, in_offset=1)
    fd.device_sync()
    rs = O.Channelizer(w.fs, chans).apply(x[1: n + 1])
    for c in range(len(chans)):
        assert counts[c] == rs[c].size
        assert_close_rms(outs[c].download(counts[c]), rs[c], xr, what=f"unaligned chan {c}")


@pytest.mark.parametrize("kernel,least", [("segment", 3), ("groups", 9)])
def test_autotune_keeps_stream_state(kernel, least):
    # tuning launches run on a chunk of the caller's data without moving the
    # stream: the chunks around it still continue the oracle's stream
    fs = 1_000_000
    chans = _chans_grouped(fs)
    x = cwgn(23, 300_000)
    g = fd.MultiRateChannelizer(fs, chans)
    g.set_kernel(kernel)
    ref = O.Channelizer(fs, chans)
    for lo, hi in [(0, 100_000), (100_000, 300_000)]:
        ys, rs = g.apply(x[lo:hi]), ref.apply(x[lo:hi])
        for c in range(len(chans)):
            assert ys[c].size == rs[c].size
            assert_close_rms(ys[c], rs[c], rms(x), what=f"chan {c} [{lo},{hi})")
        if lo == 0:
            d = fd.DeviceBuffer.from_array(x[:131_072])
            trials = g.autotune(d, 131_072, reps=3)
            timed = [t for t in trials if t["us"] > 0]
            assert len(trials) >= least and timed
            assert all(t["segment_kernel"] == (kernel == "segment") for t in trials)
            best = min(timed, key=lambda t: t["us"])
            assert g.get_tiling() == (best["segment_len"], best["block_threads"])


@pytest.mark.parametrize("kernel", ["segment", "groups"])
def test_autotune_pick_persists(kernel, tmp_path, monkeypatch):
    # the pick goes into the tuning cache file; a new channelizer with the same
    # channels, chunk length and kernel takes it from there without trials,
    # and a different chunk length (another key) runs its trials again
    cache = tmp_path / "tuning.txt"
    monkeypatch.setenv("FDSP_TUNE_CACHE", str(cache))
    fs = 1_000_000
    chans = _chans_grouped(fs)
    d = fd.DeviceBuffer.from_array(cwgn(24, 131_072))
    a = fd.MultiRateChannelizer(fs, chans)
    a.set_kernel(kernel)
    first = a.autotune(d, 131_072, reps=2)
    assert len(first) > 1 and not any(t["cached"] for t in first)
    assert cache.read_text().count("\n") == 1
    b = fd.MultiRateChannelizer(fs, chans)
    b.set_kernel(kernel)
    again = b.autotune(d, 131_072, reps=2)
    assert len(again) == 1 and again[0]["cached"]
    assert b.get_tiling() == a.get_tiling() == (again[0]["segment_len"], again[0]["block_threads"])
    assert b.get_fusion()[0] == a.get_fusion()[0]
    # the cached plan computes what the tuned one does
    x = cwgn(25, 100_000)
    for ya, yb in zip(a.apply(x), b.apply(x)):
        np.testing.assert_array_equal(yb, ya)
    other = b.autotune(d, 65_536, reps=2)
    assert len(other) > 1 and not any(t["cached"] for t in other)
    assert cache.read_text().count("\n") == 2


@pytest.mark.parametrize("seg,block", [(256, 128), (1024, 256), (2048, 512), (512, 512), (4096, 512),
                                       (fd.TILING_STRIDED, 128), (fd.TILING_CLAIM, 512)])
def test_tilings_agree(seg, block):
    fs = 20_000_000
    chans = _chans_small(fs)
    x = cwgn(10, 120_000)
    b0 = fd.MultiRateChannelizer(fs, chans)
    b0.set_kernel("groups")
    base = b0.apply(x)
    g = fd.MultiRateChannelizer(fs, chans)
    g.set_kernel("groups")
    g.set_tiling(seg, block)
    # the work split only changes where the NCO phasor recurrence restarts
    # (a few float32 rounding steps), so tilings agree to ~1e-7
    for c, (a, b) in enumerate(zip(base, g.apply(x))):
        assert a.size == b.size
        assert_close_rms(b, a, rms(x), tol=TILING_TOL, what=f"chan {c}")
