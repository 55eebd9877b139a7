"""Serving loop (fdsp_serve_*) on the GPU vs the oracle.

A stream is one channelizer channel (reference
src/transform/falcon_dsp_multi_rate_channelizer.cc:146-164), optionally with
the polar discriminator (falcon_dsp_polar_discriminator.cc:134-171) on its
output. Bounds:
- each stream's chunk results equal one channel of the oracle channelizer fed
  the same chunks: exact counts, max |err| <= 1e-4 * input RMS; tails within
  1e-3 rad of the oracle discriminator's angles (as the channelizer's own
  tail test);
- a stream's results are bit-identical alone and in a mix of other streams
  (no output's arithmetic depends on the batch it lands in);
- admission stops at the stream limit and at the device-memory budget, and
  resumes after a release.
"""
import numpy as np
import pytest

import libfalcon_dsp_amd as fd
from libfalcon_dsp_amd._native import FDSP_EBUSY, FDSP_EINVAL, FDSP_ERANGE
from libfalcon_dsp_amd.signals import cwgn, fm_signal
from libfalcon_dsp_amd.workloads import Serving, channel
from oracle import oracle as O
from tests.helpers import assert_close_rms, chunks, rms

pytestmark = pytest.mark.gpu

FS = 25_000_000
ANGLE_TOL = 1e-3


def _streams():
    return [
        channel(FS, 1_333_333, 250_000),
        channel(FS, -6_000_000, 2_500_000),
        channel(FS, 8_666_667, 12_500),
        channel(FS, -2_000_000, 250_000, tail="discriminator"),
        channel(FS, 3_000_000, 1_600_000),  # up 8, down 125
        channel(FS, -1, 100_000),
    ]


def _input(s, ch, n):
    x = cwgn(100 + s, n).astype(np.complex128)
    if ch.get("tail"):
        sig, _ = fm_signal(n, FS, [400.0, 1300.0, 2700.0], 75e3, carrier_hz=-ch["freq_shift_in_hz"])
        x = 0.05 * x + sig
    return x.astype(np.complex64)


def _oracle(ch, x, bounds):
    """Per-chunk expected outputs of one stream."""
    ref = O.Channelizer(FS, [ch])
    disc = O.Discriminator() if ch.get("tail") else None
    out = []
    for lo, hi in bounds:
        y = ref.apply(x[lo:hi])[0]
        out.append(disc.apply(y) if disc else y)
    return out


def _check(ch, got, want, x, what):
    assert got.size == want.size, f"{what}: {got.size} outputs, oracle {want.size}"
    if ch.get("tail"):
        assert got.dtype == np.float32
        d = np.angle(np.exp(1j * (got.astype(np.float64) - want)))
        assert d.size == 0 or np.max(np.abs(d)) <= ANGLE_TOL, f"{what}: angle error {np.max(np.abs(d)):.2e}"
    else:
        assert_close_rms(got, want, rms(x), what=what)


def _run_interleaved(srv, sids, xs, bounds):
    """Submit every stream's chunks round-robin without waiting, then collect."""
    tickets = [[] for _ in sids]
    k = 0
    while any(k < len(b) for b in bounds):
        for i, sid in enumerate(sids):
            if k < len(bounds[i]):
                lo, hi = bounds[i][k]
                tickets[i].append(srv.submit(sid, xs[i][lo:hi]))
        k += 1
    got = [[srv.result(t)[0] for t in ts] for ts in tickets]
    return got


def test_streams_match_oracle_interleaved():
    chans = _streams()
    n = 300_000
    sizes = [[65_536, 1, 7, 100_003], [4096, 131_072, 3], [200_000, 1999, 5], [1, 30_000, 77_777, 4096],
             [12_345, 65_536], [256 * 1024, 2, 9]]
    xs = [_input(s, ch, n) for s, ch in enumerate(chans)]
    bounds = [chunks(n, sz) for sz in sizes]
    srv = fd.Server(FS, max_streams=16, max_chunk=1 << 18, max_batch=1 << 20)
    sids = [srv.admit(ch) for ch in chans]
    assert None not in sids
    got = _run_interleaved(srv, sids, xs, bounds)
    for i, ch in enumerate(chans):
        want = _oracle(ch, xs[i], bounds[i])
        for k, (g, w) in enumerate(zip(got[i], want)):
            _check(ch, g, w, xs[i], f"stream {i} chunk {k} {bounds[i][k]}")
    info = srv.info()
    assert info["chunks"] == sum(len(b) for b in bounds)
    assert info["batches"] < info["chunks"]  # chunks of several streams shared launches
    srv.close()


def test_alone_equals_in_the_mix():
    target = channel(FS, -3_141_593, 250_000)
    others = [channel(FS, int(-9e6 + 0.9e6 * k), [12_500, 100_000, 1_000_000, 2_500_000][k % 4]) for k in range(20)]
    others[5]["tail"] = "discriminator"
    n = 400_000
    x = _input(0, target, n)
    b = chunks(n, [50_000, 1, 70_001])
    alone_srv = fd.Server(FS, max_streams=4)
    sid = alone_srv.admit(target)
    alone = [alone_srv.result(alone_srv.submit(sid, x[lo:hi]))[0] for lo, hi in b]
    alone_srv.close()

    mix_srv = fd.Server(FS, max_streams=32, max_batch=1 << 19)
    sids = [mix_srv.admit(c) for c in [target] + others]
    xs = [x] + [_input(k + 1, c, n) for k, c in enumerate(others)]
    bounds = [b] + [chunks(n, [30_000 + 1000 * k, 5]) for k in range(len(others))]
    mixed = _run_interleaved(mix_srv, sids, xs, bounds)[0]
    assert mix_srv.info()["batches"] < sum(len(q) for q in bounds)
    mix_srv.close()
    assert len(alone) == len(mixed)
    for k, (a, m) in enumerate(zip(alone, mixed)):
        np.testing.assert_array_equal(a, m, err_msg=f"chunk {k}")
    want = _oracle(target, x, b)
    for k, (a, w) in enumerate(zip(alone, want)):
        _check(target, a, w, x, f"chunk {k}")


def test_admission_by_budget_and_stream_limit():
    ch = channel(FS, 1_000_000, 250_000)
    # the batch buffers' size depends on the configuration: probe with the same one
    probe = fd.Server(FS, max_streams=8, max_chunk=1 << 16, max_batch=1 << 16, hbm_budget=0)
    fixed = probe.info()["hbm_used"]
    s0 = probe.admit(ch)
    per_stream = probe.info()["hbm_used"] - fixed
    assert per_stream > 0
    probe.release(s0)
    assert probe.info()["hbm_used"] == fixed
    probe.close()

    srv = fd.Server(FS, max_streams=8, max_chunk=1 << 16, max_batch=1 << 16, hbm_budget=fixed + 2 * per_stream)
    a, b = srv.admit(ch), srv.admit(ch)
    assert a is not None and b is not None
    assert srv.admit(ch) is None  # over the budget
    srv.release(a)
    c = srv.admit(ch)
    assert c is not None
    x = cwgn(7, 10_000)
    y, res = srv.result(srv.submit(c, x))
    assert res.status == 0 and y.size == O.Channelizer(FS, [ch]).apply(x)[0].size
    srv.close()

    lim = fd.Server(FS, max_streams=3, max_chunk=1 << 16, max_batch=1 << 16)
    ids = [lim.admit(ch) for _ in range(3)]
    assert None not in ids and lim.admit(ch) is None
    lim.release(ids[1])
    assert lim.admit(ch) is not None
    lim.close()

    with pytest.raises(fd.FdspError):  # batch buffers alone above the budget
        fd.Server(FS, max_chunk=1 << 16, max_batch=1 << 16, hbm_budget=1 << 20)


def test_error_codes():
    ch = channel(FS, 2_000_000, 100_000)
    srv = fd.Server(FS, max_streams=4, max_chunk=1 << 16, max_batch=1 << 17)
    sid = srv.admit(ch)
    with pytest.raises(fd.FdspError) as e:
        srv.submit(sid + 1, cwgn(1, 100))
    assert e.value.code == FDSP_EINVAL
    with pytest.raises(fd.FdspError) as e:
        srv.submit(sid, cwgn(1, (1 << 16) + 1))
    assert e.value.code == FDSP_ERANGE
    with pytest.raises(fd.FdspError) as e:
        srv.poll(123456789)
    assert e.value.code == FDSP_EINVAL
    # a stream with requests queued cannot be released
    x = cwgn(2, 1 << 16)
    ts = [srv.submit(sid, x) for _ in range(64)]
    with pytest.raises(fd.FdspError) as e:
        srv.release(sid)
    assert e.value.code == FDSP_EBUSY
    for t in ts:
        srv.result(t)
    srv.release(sid)
    srv.close()
    with pytest.raises(fd.FdspError) as e:
        fd.Server(FS, max_chunk=1 << 18, max_batch=1 << 17)
    assert e.value.code == FDSP_EINVAL


def test_serving_workload_slice():
    w = Serving(n_requests=240, max_streams=40)
    reqs = w.requests()
    srv = fd.Server(w.fs, max_streams=w.max_streams, max_chunk=w.chunk_hi, max_batch=1 << 21,
                    hbm_budget=w.hbm_budget)
    sids, pos, fed, tickets = {}, {}, {}, []
    for r in reqs:
        if r.stream not in sids:
            sids[r.stream] = srv.admit(w.stream_channel(r.stream))
            pos[r.stream], fed[r.stream] = 0, []
        x = w.stream_input(r.stream, pos[r.stream], r.chunk)
        fed[r.stream].append(x)
        pos[r.stream] += r.chunk
        tickets.append((r.stream, srv.submit(sids[r.stream], x)))
    got = {s: [] for s in sids}
    lat = []
    for s, t in tickets:
        y, res = srv.result(t)
        got[s].append(y)
        lat.append(res.done_ns - res.submit_ns)
    assert min(lat) > 0
    checked = 0
    for s in sorted(sids)[:12]:
        ch = w.stream_channel(s)
        x = np.concatenate(fed[s])
        b, i = [], 0
        for part in fed[s]:
            b.append((i, i + part.size))
            i += part.size
        for k, (g, want) in enumerate(zip(got[s], _oracle(ch, x, b))):
            _check(ch, g, want, x, f"stream {s} chunk {k}")
            checked += 1
    assert checked > 12
    srv.close()


def test_request_timestamps_and_state_pool():
    # every result carries its stages in order (submit, taken into a batch,
    # copy-in issued, GPU work seen done, delivered); admission takes slices
    # of the pool allocated at create, and a released slice is reused
    ch = channel(FS, 1_000_000, 250_000)
    srv = fd.Server(FS, max_streams=4, max_chunk=1 << 16, max_batch=1 << 17, state_bytes=1 << 20)
    info = srv.info()
    assert info["state_pool"] == 1 << 20
    sids = [srv.admit(ch) for _ in range(3)]
    used = srv.info()["hbm_used"]
    x = cwgn(9, 50_000)
    ts = [srv.submit(s, x) for s in sids for _ in range(3)]
    for t in ts:
        y, r = srv.result(t)
        assert r.status == 0
        assert r.submit_ns <= r.pack_ns <= r.issue_ns <= r.gpu_ns <= r.done_ns
    srv.release(sids[0])
    assert srv.info()["hbm_used"] < used
    assert srv.admit(ch) is not None
    assert srv.info()["hbm_used"] == used
    srv.close()
