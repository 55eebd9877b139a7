"""Segment-major FP32 channelizer kernel (FDSP_KERNEL_SEGMENT_F32), the
batched default, against the oracle (reference
src/transform/falcon_dsp_multi_rate_channelizer.cc:146-164 over
falcon_dsp_freq_shift.cc:164-192 and falcon_dsp_polyphase_resampler.cc:256-319).

Every output's window, taps and NCO index are exact whatever the segment
length or the chunking of the stream; the NCO phasor advances by an FP32
recurrence from an FP64 seed at each wave's first sub-tile of a segment, so
segment lengths and chunkings agree to a few FP32 roundings (F32_TOL), and
the same chunking gives bit-identical results eagerly and in graph mode.
The arithmetic is FP32
throughout (FP32 matrix cores and vector ALUs), as the reference's
multiply-accumulates are (polyphase_resampler.cc:288,298): on the 32-channel
workload its error against the oracle stays below F32_TOL x input RMS, a
bound the opt-in bf16x2 form (bf16 hi/lo planes, about 16 significant bits)
does not meet.
"""
import numpy as np
import pytest

import libfalcon_dsp_amd as fd
from libfalcon_dsp_amd.signals import cwgn
from libfalcon_dsp_amd.workloads import Chan32, channel
from oracle import oracle as O
from tests.helpers import assert_close_rms, chunks, max_err, rms

pytestmark = pytest.mark.gpu

# max |err| against the oracle relative to the input RMS that FP32 arithmetic
# meets on the 32-channel workload (measured about 1e-7) and the bf16x2 form
# does not (measured about 1e-5)
F32_TOL = 1e-6


def _mixed(fs=20_000_000):
    # MFMA groups (down 8 x3, down 80 x2), VALU channels of up 1 (down 20,
    # a 48-tap window of 64), up > 1 (up 3 down 20; up 2 down 25) and an
    # identity shift
    return ([channel(fs, f, 2_500_000) for f in (1_234_567, -3_000_001, 6_000_000)]
            + [channel(fs, f, 250_000) for f in (-8_000_000, 77)]
            + [channel(fs, 781_283, 1_000_000), channel(fs, 0, 12_500, taps_per_up=48),
               channel(fs, -1, 3_000_000), channel(fs, 2_000_000, 1_600_000)])


def _run(fs, chans, x, sizes, seg=None, kernel="segment"):
    g = fd.MultiRateChannelizer(fs, chans)
    g.set_kernel(kernel)
    if seg:
        g.set_tiling(seg, 256)
    parts = [[] for _ in chans]
    for lo, hi in chunks(x.size, sizes):
        for c, y in enumerate(g.apply(x[lo:hi])):
            parts[c].append(y)
    return g, [np.concatenate(p) for p in parts]


def test_segment_kernel_is_the_batched_default():
    w = Chan32(n=1 << 16)
    g = fd.MultiRateChannelizer(w.fs, w.channels())
    info = g.kernel_info()
    # down 8 (7 channels), 80 (5), 200 (7), 800 (7), 1600 (5) on matrix cores,
    # the one channel at down 20 on the vector ALUs
    assert info["kernel"] == "chan_seg_kernel" and info["mfma_groups"] == 5 and info["groups"] == 6
    assert fd.MultiRateChannelizer(w.fs, w.channels(), arith="bf16x2").kernel_info()["kind"] == 2
    # a window above 64 taps: the channel-group kernel, still FP32
    wide = fd.MultiRateChannelizer(20_000_000, [channel(20_000_000, 5, 250_000, taps_per_up=80)])
    assert wide.kernel_info()["kernel"] == "chan_groups_kernel"
    with pytest.raises(fd.FdspError):
        wide.set_kernel("segment")


def test_mixed_channel_set_against_oracle():
    fs = 20_000_000
    chans = _mixed(fs)
    x = cwgn(31, 300_000)
    g = fd.MultiRateChannelizer(fs, chans)
    ref = O.Channelizer(fs, chans)
    xr = rms(x)
    for lo, hi in chunks(x.size, [65_536, 1, 3, 100_003, 17, 2048, 4095]):
        print(f"segment kernel chunk [{lo},{hi})", flush=True)
        ys, rs = g.apply(x[lo:hi]), ref.apply(x[lo:hi])
        for c, (y, r) in enumerate(zip(ys, rs)):
            assert y.size == r.size, f"chan {c} [{lo},{hi})"
            assert_close_rms(y, r, xr, tol=F32_TOL, what=f"chan {c} [{lo},{hi})")
            k, ph, xo, it, ot = g.state(c)
            rk, rph, rxo = ref.state(c)
            assert (ph, xo) == (rph, rxo) and it == hi
            if chans[c]["freq_shift_in_hz"]:
                assert k == rk


@pytest.mark.parametrize("seg", [1024, 1280, 3072, 3840, 4096])
def test_segment_lengths_agree(seg):
    fs = 20_000_000
    chans = _mixed(fs)
    x = cwgn(32, 200_000)
    _, base = _run(fs, chans, x, [x.size], seg=2048)
    _, other = _run(fs, chans, x, [x.size], seg=seg)
    for c in range(len(chans)):
        assert other[c].size == base[c].size
        assert_close_rms(other[c], base[c], rms(x), tol=F32_TOL, what=f"chan {c}")


def test_segment_length_limits():
    g = fd.MultiRateChannelizer(20_000_000, _mixed())
    for bad in (512, 768, 1000, 2560 + 128, 4352):
        with pytest.raises(fd.FdspError):
            g.set_tiling(bad, 256)
    assert g.get_tiling()[0] == 2560  # the default
    g.set_tiling(1280, 256)
    assert g.get_tiling()[0] == 1280


def test_chunkings_agree():
    # the NCO index and the window of every output are exact whatever the
    # chunk boundaries, so any split of the stream gives the same samples up
    # to the phasor recurrence's FP32 rounding
    fs = 1_000_000
    chans = [channel(fs, 23_000, 50_000), channel(fs, -7, 50_000), channel(fs, 333_333, 250_000),
             channel(fs, 100_001, 12_500, taps_per_up=48)]
    x = cwgn(33, 900_000)
    _, whole = _run(fs, chans, x, [x.size])
    _, split = _run(fs, chans, x, [434_780, 5, 1, 131_071, 3])
    for c in range(len(chans)):
        assert split[c].size == whole[c].size
        assert_close_rms(split[c], whole[c], rms(x), tol=F32_TOL, what=f"chan {c}")


def test_fp32_faithful_on_the_chan32_slice():
    # the 32-channel workload (2^20 samples of it, in 2^18-sample chunks):
    # the FP32 kernels stay within F32_TOL x RMS of the oracle, the opt-in
    # bf16x2 form does not
    w = Chan32(n=1 << 20)
    chans = w.channels()
    x = w.input()
    xr = rms(x)
    errs = {}
    for name, kw in (("segment_f32", {}), ("groups_f32", {"kernel": "groups"}), ("bf16x2", None)):
        if kw is None:
            g = fd.MultiRateChannelizer(w.fs, chans, arith="bf16x2")
        else:
            g = fd.MultiRateChannelizer(w.fs, chans)
            g.set_kernel(kw.get("kernel", "segment"))
        ref = O.Channelizer(w.fs, chans)
        worst = 0.0
        for lo, hi in chunks(x.size, [1 << 18]):
            ys, rs = g.apply(x[lo:hi]), ref.apply(x[lo:hi])
            for c in range(len(chans)):
                assert ys[c].size == rs[c].size
                worst = max(worst, max_err(ys[c], rs[c]) / xr)
        errs[name] = worst
    print("max |err| / input RMS against the oracle:", {k: f"{v:.3e}" for k, v in errs.items()}, flush=True)
    assert errs["segment_f32"] <= F32_TOL
    assert errs["groups_f32"] <= F32_TOL
    assert F32_TOL < errs["bf16x2"] <= 1e-4


def test_graph_replays_match_eager():
    # graph mode (device-resident position) equals eager calls bit for bit
    fs = 20_000_000
    chans = _mixed(fs)
    chunk = 50_000
    x = cwgn(34, chunk * 4)
    g = fd.MultiRateChannelizer(fs, chans)
    d_in = fd.DeviceBuffer(chunk)
    outs = [fd.DeviceBuffer(m) for m in g.max_out(chunk)]
    g.graph_create(chunk, d_in, outs)
    e = fd.MultiRateChannelizer(fs, chans)
    s = fd.Stream()
    for k in range(4):
        d_in.upload(x[k * chunk:(k + 1) * chunk], s)
        g.graph_launch(s)
        s.sync()
        counts = g.graph_counts()
        want = e.apply(x[k * chunk:(k + 1) * chunk])
        for c in range(len(chans)):
            assert counts[c] == want[c].size
            np.testing.assert_array_equal(outs[c].download(counts[c]), want[c], err_msg=f"replay {k} chan {c}")


def _fm_mix(fs=25_000_000):
    # FM streams with discriminator tails: down 100, up 2 down 125 (at most 63
    # outputs per segment, one batch) and down 5 with a 64-tap window (several
    # batches of 63 per segment, as in FM demodulation), beside plain channels
    # that share an MFMA group
    return [channel(fs, 2_000_000, 250_000, tail="discriminator"),
            channel(fs, -3_500_000, 400_000, tail="discriminator"),
            channel(fs, 1_000_000, 250_000), channel(fs, -6_000_000, 250_000), channel(fs, 77, 2_500_000),
            channel(fs, 2_000_000, 5_000_000, taps_per_up=64, tail="discriminator")]


def _fm_input(fs, n, seed):
    from libfalcon_dsp_amd.signals import fm_signal
    a, _ = fm_signal(n, fs, [400.0, 1100.0, 2900.0], 75e3, carrier_hz=-2_000_000)
    b, _ = fm_signal(n, fs, [700.0, 2300.0], 75e3, carrier_hz=3_500_000)
    return (a + b + 0.01 * cwgn(seed, n)).astype(np.complex64)


def _check_fm(got, ref_ch, ref_disc, c, what):
    want = ref_disc.apply(ref_ch) if ref_disc is not None else ref_ch
    assert got.dtype == np.float32 and got.size == want.size, f"{what} chan {c}: {got.size} vs {want.size}"
    if got.size:
        d = np.angle(np.exp(1j * (got.astype(np.float64) - want)))
        assert np.max(np.abs(d)) <= 1e-3, f"{what} chan {c}: max angle error {np.max(np.abs(d)):.2e}"


def test_fused_discriminator_against_oracle():
    # the angles come out of the channelizer kernel itself: chunks of every
    # size, including ones that emit no output for a tail channel (fewer
    # inputs than down) and single samples
    fs = 25_000_000
    chans = _fm_mix(fs)
    x = _fm_input(fs, 400_000, 35)
    g = fd.MultiRateChannelizer(fs, chans)
    assert g.kernel_info()["kernel"] == "chan_seg_kernel"
    ref = O.Channelizer(fs, chans)
    discs = [O.Discriminator() if ch.get("tail") else None for ch in chans]
    xr = rms(x)
    for lo, hi in chunks(x.size, [100_000, 3, 1, 50, 77_777, 99, 2048, 4097]):
        print(f"fused discriminator chunk [{lo},{hi})", flush=True)
        ys, rs = g.apply(x[lo:hi]), ref.apply(x[lo:hi])
        for c, ch in enumerate(chans):
            if ch.get("tail"):
                _check_fm(ys[c], rs[c], discs[c], c, f"[{lo},{hi})")
            else:
                assert ys[c].size == rs[c].size
                assert_close_rms(ys[c], rs[c], xr, tol=F32_TOL, what=f"chan {c} [{lo},{hi})")


def test_fused_discriminator_graph_replays():
    # graph mode carries the discriminator's last sample on the device
    fs = 25_000_000
    chans = _fm_mix(fs)
    chunk = 65_536
    x = _fm_input(fs, chunk * 5, 36)
    g = fd.MultiRateChannelizer(fs, chans)
    d_in = fd.DeviceBuffer(chunk)
    outs = [fd.DeviceBuffer(m, np.float32 if ch.get("tail") else np.complex64)
            for m, ch in zip(g.max_out(chunk), chans)]
    g.graph_create(chunk, d_in, outs)
    ref = O.Channelizer(fs, chans)
    discs = [O.Discriminator() if ch.get("tail") else None for ch in chans]
    s = fd.Stream()
    for k in range(5):
        d_in.upload(x[k * chunk:(k + 1) * chunk], s)
        g.graph_launch(s)
        s.sync()
        counts = g.graph_counts()
        rs = ref.apply(x[k * chunk:(k + 1) * chunk])
        for c, ch in enumerate(chans):
            got = outs[c].download(counts[c])
            if ch.get("tail"):
                _check_fm(got, rs[c], discs[c], c, f"replay {k}")
            else:
                assert counts[c] == rs[c].size
                assert_close_rms(got, rs[c], rms(x), tol=F32_TOL, what=f"replay {k} chan {c}")
