"""Benchmark: aggregate output Msps of the 32-channel multi-rate channelizer
(2^24 complex64 samples at 20 Msps, 32 channels, one fused launch per step,
replayed as a HIP graph) with p99 chunk latency, roofline and the reference
CPU channelizer timed beside it.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--no-cpu] [--no-plain] [--no-serving] [--no-others]

A step is one pass of the batched channelizer over the whole 2^24-sample
input, resident in HBM before timing starts. The same line carries the
serving workload (p50/p99 request latency, aggregate Msps) and the
resampler, multi-shift NCO and FIR-bank workloads. Prints ONE JSON line on
stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(fs, chans, x, n_single, n_par):
    """The reference's own CPU channelizer (oracle/_ref) when it was built,
    else the oracle restatement: single thread over n_single samples, then
    channel-parallel over n_par samples with one object per thread."""
    from oracle import oracle as O
    kind = "reference" if O.ref_available() else "port"
    make = O.RefChannelizer if kind == "reference" else O.Channelizer

    r = make(fs, chans)
    t0 = time.perf_counter()
    ys = r.apply(x[:n_single])
    t1 = time.perf_counter()
    single = sum(y.size for y in ys) / (t1 - t0) / 1e6

    threads = cpu_threads()
    groups = [chans[i::threads] for i in range(threads)]
    groups = [g for g in groups if g]
    objs = [make(fs, g) for g in groups]
    totals = [0] * len(objs)

    def work(i):
        totals[i] = sum(y.size for y in objs[i].apply(x[:n_par]))

    ths = [threading.Thread(target=work, args=(i,)) for i in range(len(objs))]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    t1 = time.perf_counter()
    par = sum(totals) / (t1 - t0) / 1e6
    return {
        "value": round(par, 3), "unit": "Msps", "cores": len(objs), "kind": kind,
        "sample": f"first {n_par} input samples (all 32 channels), channel-parallel, "
                  f"{len(objs)} threads, one reference channelizer object per thread",
        "single_thread_value": round(single, 3),
        "single_thread_sample": f"first {n_single} input samples, 1 thread",
        "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
    }


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the 32-channel workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes, each scaled by the
    same counter's reading for a device copy of known size; (None, None)
    when no such measurement is committed for that kernel."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_chan32_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return int(t["traffic_bytes_per_dispatch"]), os.path.relpath(path, ROOT)
    except (OSError, KeyError, ValueError):
        return None, None


def chan32_graph_rate(fd, w, chans, x_dev, stream, steps, warmup, arith, kernel=None, tune=True):
    """Aggregate output Msps of one channelizer form over the 32-channel
    workload, replayed as a HIP graph, and its kernel's average time from
    back-to-back eager launches (events on the kernel's stream)."""
    n = w.n
    g = fd.MultiRateChannelizer(w.fs, chans, batched=True, arith=arith)
    if kernel:
        g.set_kernel(kernel)
    trials = g.autotune(x_dev, n, reps=5) if tune else None
    outs = [fd.DeviceBuffer(m) for m in g.max_out(n)]
    g.graph_create(n, x_dev, outs)
    for _ in range(warmup):
        g.graph_launch(stream)
    stream.sync()
    e0, e1 = fd.Event(), fd.Event()
    total = 0
    e0.record(stream)
    for _ in range(steps):
        g.graph_launch(stream)
        total += sum(g.graph_counts())
    e1.record(stream)
    e1.sync()
    ms = e0.elapsed_ms(e1)
    g.apply_device(x_dev, n, outs, stream)
    stream.sync()
    k0, k1 = fd.Event(), fd.Event()
    k0.record(stream)
    for _ in range(steps):
        g.apply_device(x_dev, n, outs, stream)
    k1.record(stream)
    k1.sync()
    k_us = 1e3 * k0.elapsed_ms(k1) / steps
    seg, block = g.get_tiling()
    return {"value": round(total / (ms / 1e3) / 1e6, 1), "unit": "Msps", "ms_per_step": round(ms / steps, 5),
            "kernel": g.kernel_info()["kernel"], "kernel_avg_us": round(k_us, 2),
            "tiling": [seg, block], "autotune_trials": trials}


def pct(v, q):
    v = sorted(v)
    return v[min(len(v) - 1, int(round(q * (len(v) - 1))))]


def timed_launches(fd, stream, launch, reps):
    """Average ms of `reps` back-to-back launches between one event pair on `stream`."""
    launch()
    stream.sync()
    e0, e1 = fd.Event(), fd.Event()
    e0.record(stream)
    for _ in range(reps):
        launch()
    e1.record(stream)
    e1.sync()
    return e0.elapsed_ms(e1) / reps


def iir_cpu_baseline():
    """The reference's IIR filter (oracle/_ref; else the oracle restatement)
    on the iir_order4 filter over the first 2^20 of its samples, one thread
    (one stream's recurrence is sequential on the CPU too)."""
    from libfalcon_dsp_amd.workloads import IirOrder4
    from oracle import oracle as O
    w = IirOrder4()
    b, a = w.coeffs()
    n = 1 << 20
    x = w.input()[:n]
    kind = "reference" if O.ref_available() else "port"
    f = (O.RefIir if kind == "reference" else O.Iir)(b, a)
    t0 = time.perf_counter()
    f.apply(x)
    dt = time.perf_counter() - t0
    return {"value": round(n / dt / 1e6, 2), "unit": "output Msps", "cores": 1, "kind": kind,
            "sample": "first 2^20 samples of iir_order4, one thread"}


def other_cpu_baselines():
    """The reference's own CPU classes (oracle/_ref; else the oracle
    restatement) on bounded samples of C1, C2 and C4, in each line's unit:
    one object per thread, as the reference's objects are single-threaded."""
    from libfalcon_dsp_amd.workloads import FirBank256, Nco16, Resampler1ch
    from oracle import oracle as O
    kind = "reference" if O.ref_available() else "port"
    res = {}

    w1 = Resampler1ch()
    x1 = w1.input()
    r = (O.RefResampler if kind == "reference" else O.Resampler)(w1.up, w1.down, w1.taps())
    n_out, t0 = 0, time.perf_counter()
    for k in range(4):
        n_out += r.apply(x1[k * w1.chunk:(k + 1) * w1.chunk]).size
    dt = time.perf_counter() - t0
    res["resampler_1ch"] = {"value": round(n_out / dt / 1e6, 2), "unit": "output Msps", "cores": 1, "kind": kind,
                            "sample": "first 4 of the 16 chunks of 2^16, one stream"}

    w2 = Nco16()
    x2 = w2.input()[:1 << 20]
    make = O.RefFreqShift if kind == "reference" else O.FreqShift
    objs = [make(w2.fs, f) for f in w2.shifts()]
    t0 = time.perf_counter()
    objs[0].apply(x2)
    single = x2.size / (time.perf_counter() - t0) / 1e6
    objs = [make(w2.fs, f) for f in w2.shifts()]
    ths = [threading.Thread(target=o.apply, args=(x2,)) for o in objs]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    dt = time.perf_counter() - t0
    res["nco_16shift"] = {"value": round(16 * x2.size / dt / 1e6, 1), "unit": "output Msps",
                          "cores": len(ths), "kind": kind, "single_thread_value": round(single, 2),
                          "sample": "first 2^20 inputs, the 16 shifts one object per thread"}

    w4 = FirBank256()
    taps = w4.taps()
    n4, nch = 1 << 16, 16
    make = O.RefFir if kind == "reference" else O.Fir
    objs = [make(taps[c].astype(np.complex64)) for c in range(nch)]
    xs = [w4.input(c)[:n4] for c in range(nch)]
    ths = [threading.Thread(target=objs[c].apply, args=(xs[c],)) for c in range(nch)]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    dt = time.perf_counter() - t0
    res["fir_bank_256"] = {"value": round(4.0 * w4.n_taps * n4 * nch / dt / 1e12, 4), "unit": "TFLOP/s",
                           "cores": len(ths), "kind": kind,
                           "sample": "2^16 samples of 16 of the 256 channels, one object per thread (4 flops per "
                                     "real tap per output, as the GPU line counts them)"}
    return res


def bench_other_configs(fd, stream, steps):
    """The other workloads, each with data resident in HBM: throughput and
    roofline fraction of one pass."""
    import ctypes as C

    from libfalcon_dsp_amd._native import check
    from libfalcon_dsp_amd.workloads import FirBank256, Nco16, Resampler1ch
    out = {}

    # resampler_1ch: 16 chunks of 2^16 through the streaming resampler
    w1 = Resampler1ch()
    x1 = w1.input()
    r = fd.PolyphaseResampler(w1.up, w1.down, w1.taps())
    d1 = fd.DeviceBuffer.from_array(x1)
    o1 = fd.DeviceBuffer(w1.n * w1.up // w1.down + 16)
    cnt = [0]

    def pass1_each():
        # one launch per chunk, as a per-chunk streaming caller makes them
        cnt[0] = 0
        for k in range(w1.n // w1.chunk):
            cnt[0] += r.apply_device(d1, w1.chunk, o1, stream, in_offset=k * w1.chunk)

    def pass1():
        # the 16 chunks in one launch (fdsp_resampler_apply_chunks_device); the
        # stream continues across passes, counts differ by at most one per chunk
        cnt[0] = sum(r.apply_chunks_device(d1, w1.chunk, w1.n // w1.chunk, o1, stream))

    ms1e = timed_launches(fd, stream, pass1_each, steps)
    ms1 = timed_launches(fd, stream, pass1, steps)
    b1 = 8 * w1.n + 8 * cnt[0]
    out["resampler_1ch"] = {
        "value": round(cnt[0] / (ms1 / 1e3) / 1e6, 1), "unit": "output Msps", "ms_per_pass": round(ms1, 4),
        "outputs": cnt[0], "roofline": {"bound": "hbm", "achieved": round(b1 / (ms1 / 1e3) / 1e9, 1),
                                         "peak": HBM_PEAK_GBS, "unit": "GB/s"},
        "per_chunk_launches": {"value": round(cnt[0] / (ms1e / 1e3) / 1e6, 1), "unit": "output Msps",
                               "ms_per_pass": round(ms1e, 4)},
        "note": "2^20 inputs as 16 chunks of 2^16 in one launch; per_chunk_launches makes 16 launches"}
    out["resampler_1ch"]["roofline"]["frac"] = round(out["resampler_1ch"]["roofline"]["achieved"] / HBM_PEAK_GBS, 4)
    log(f"resampler_1ch: {ms1:.3f} ms per 16-chunk launch ({ms1e:.3f} ms as 16 launches), "
        f"{out['resampler_1ch']['value']} Msps")
    del d1, o1, r

    # nco_16shift: 64 chunks of 2^18, 16 shifted copies each
    w2 = Nco16()
    x2 = w2.input()
    f = fd.FreqShift(w2.fs, w2.shifts())
    d2 = fd.DeviceBuffer.from_array(x2)
    o2 = fd.DeviceBuffer(16 * w2.chunk)
    o2all = fd.DeviceBuffer(16 * w2.n)  # 2 GiB: every output sample lands in HBM

    def pass2_each():
        # one launch per chunk into one 2^18 x 16 buffer, as a per-chunk caller makes them
        for k in range(w2.n // w2.chunk):
            f.apply_device(d2, w2.chunk, o2, w2.chunk, stream, in_offset=k * w2.chunk)

    def pass2():
        # the 64 chunks in one launch (fdsp_freq_shift_apply_chunks_device), distinct outputs
        f.apply_chunks_device(d2, w2.chunk, w2.n // w2.chunk, o2all, w2.n, stream)

    ms2e = timed_launches(fd, stream, pass2_each, max(2, steps // 4))
    ms2 = timed_launches(fd, stream, pass2, max(2, steps // 4))
    del o2all
    b2 = 8 * w2.n + 8 * 16 * w2.n
    out["nco_16shift"] = {
        "value": round(16 * w2.n / (ms2 / 1e3) / 1e6, 1), "unit": "output Msps", "ms_per_pass": round(ms2, 4),
        "roofline": {"bound": "hbm", "achieved": round(b2 / (ms2 / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(b2 / (ms2 / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "per_chunk_launches": {"value": round(16 * w2.n / (ms2e / 1e3) / 1e6, 1), "unit": "output Msps",
                               "ms_per_pass": round(ms2e, 4),
                               "note": "64 launches rewriting one 32 MiB output buffer (cache resident)"},
        "note": "2^24 inputs as 64 chunks of 2^18 in one launch, 16 shifts into a 2 GiB output, "
                "136 B per input sample"}
    log(f"nco_16shift: {ms2:.3f} ms per launch ({ms2e:.3f} ms as 64 launches), "
        f"{out['nco_16shift']['roofline']['achieved']} GB/s")
    # error of one fresh chunk against the FP64 closed form x * e^{j w (n mod R)} (the reference's
    # rollover index), and against continuous phase x * e^{j w n}, relative to the input RMS
    f0 = fd.FreqShift(w2.fs, w2.shifts())
    f0.apply_device(d2, w2.chunk, o2, w2.chunk, stream)
    stream.sync()
    y = o2.download(16 * w2.chunk).reshape(16, w2.chunk).astype(np.complex128)
    xc = x2[:w2.chunk].astype(np.complex128)
    rms = np.sqrt(np.mean(np.abs(xc) ** 2))
    n_idx = np.arange(w2.chunk, dtype=np.float64)
    e_roll = e_cont = 0.0
    for c, f_hz in enumerate(w2.shifts()):
        r_c, om = fd.get_freq_shift_params(w2.fs, f_hz)
        e_roll = max(e_roll, float(np.max(np.abs(y[c] - xc * np.exp(1j * om * (n_idx % r_c))))))
        e_cont = max(e_cont, float(np.max(np.abs(y[c] - xc * np.exp(1j * om * n_idx)))))
    out["nco_16shift"]["fp64_error"] = {"vs_rollover_index": round(e_roll / rms, 9),
                                        "vs_continuous_phase": round(e_cont / rms, 6),
                                        "unit": "max abs error / input RMS, first 2^18-sample chunk, 16 shifts"}
    log(f"nco_16shift error vs FP64: {e_roll / rms:.2e} (rollover index), {e_cont / rms:.2e} (continuous phase)")
    del f0, x2, y
    del d2, o2

    # FM mono demodulation, 2^24 samples at 1.14 Msps in 2^18-sample chunks
    from libfalcon_dsp_amd.workloads import FmDemod
    w5 = FmDemod()
    iso, mono = w5.taps()
    fm = fd.FmDemodulator(w5.fs, w5.offset, iso, mono)
    x5 = w5.input()
    d5 = fd.DeviceBuffer.from_array(x5)
    cap5 = fm.out_count(w5.chunk) + 8
    o5 = fd.DeviceBuffer(cap5, np.int16)
    got5 = C.c_size_t()

    def pass5():
        for k in range(w5.n // w5.chunk):
            check(fd.lib().fdsp_fm_demod_apply_device(fm.h, d5.at(k * w5.chunk), w5.chunk, o5.ptr, None, cap5,
                                                      C.byref(got5), stream.h))

    ms5 = timed_launches(fd, stream, pass5, max(2, steps // 4))
    b5 = 8 * w5.n + 2 * (w5.n // 25)
    out["fm_demod"] = {
        "value": round(w5.n / (ms5 / 1e3) / 1e6, 1), "unit": "input Msps", "ms_per_pass": round(ms5, 4),
        "realtime_factor": round((w5.n / w5.fs) / (ms5 / 1e3), 1),
        "roofline": {"bound": "hbm", "achieved": round(b5 / (ms5 / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(b5 / (ms5 / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "note": "2^24 cf32 samples at 1.14 Msps in 64 chunks of 2^18; channelizer pass with a discriminator "
                "tail, de-emphasis, audio decimator, rescale; 8 B in per input sample, int16 audio out"}
    log(f"fm_demod: {ms5:.3f} ms per pass, {out['fm_demod']['value']} Msps in, "
        f"{out['fm_demod']['realtime_factor']}x real time")
    del d5, o5, fm, x5

    # the same shifts on complex int16 samples in and out (4 B each way)
    x3 = w2.input_ci16()
    d3 = fd.DeviceBuffer.from_array(x3.view(fd.CI16).ravel())
    o3 = fd.DeviceBuffer(16 * w2.chunk, fd.CI16)

    def pass3():
        for k in range(w2.n // w2.chunk):
            f.apply_ci16_device(d3, w2.chunk, o3, w2.chunk, stream, in_offset=k * w2.chunk)

    ms3 = timed_launches(fd, stream, pass3, max(2, steps // 4))
    b3 = 4 * w2.n + 4 * 16 * w2.n
    out["nco_16shift_ci16"] = {
        "value": round(16 * w2.n / (ms3 / 1e3) / 1e6, 1), "unit": "output Msps", "ms_per_pass": round(ms3, 4),
        "roofline": {"bound": "hbm", "achieved": round(b3 / (ms3 / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(b3 / (ms3 / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "note": "as nco_16shift with complex int16 in and out (68 B per input sample); the reference's "
                "per-sample phasors from a host-built table, as the reference computes int16 outputs"}
    log(f"nco_16shift_ci16: {ms3:.3f} ms per pass, {out['nco_16shift_ci16']['roofline']['achieved']} GB/s")
    del d3, o3, f, x3

    # iir_order4: 2^24 complex samples through a 4th-order Butterworth, 16 chunks of 2^20,
    # in the default mode (float outputs within rounding: this filter's segments do not
    # merge bit for bit) and in exact mode (the reference's outputs bit for bit)
    from libfalcon_dsp_amd.workloads import IirOrder4
    w6 = IirOrder4()
    bq, aq = w6.coeffs()
    d6 = fd.DeviceBuffer.from_array(w6.input())
    o6 = fd.DeviceBuffer(w6.n)
    for key, exact in (("iir_order4", None), ("iir_order4_exact", True)):
        g6 = fd.IirFilter(bq, aq)
        g6.set_exact(exact)

        def pass6():
            for k in range(w6.n // w6.chunk):
                check(fd.lib().fdsp_iir_apply_device(g6.h, d6.at(k * w6.chunk), w6.chunk, o6.at(k * w6.chunk),
                                                     stream.h))

        ms6 = timed_launches(fd, stream, pass6, 2 if exact else max(2, steps // 4))
        b6 = 16 * w6.n
        out[key] = {
            "value": round(w6.n / (ms6 / 1e3) / 1e6, 1), "unit": "output Msps", "ms_per_pass": round(ms6, 4),
            "plan": g6.plan, "warm": g6.warm,
            "roofline": {"bound": "hbm", "achieved": round(b6 / (ms6 / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(b6 / (ms6 / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
            "note": "2^24 cf32 samples in 16 chunks of 2^20, 4th-order Butterworth (5 b, 5 a taps), 16 B per "
                    "sample; " + ("bit-identical to the reference's loop, which for this filter is one sequential "
                                  "recurrence per stream" if exact else "parallel segments warmed up 160 samples, "
                                  "float outputs within rounding of the reference's (tested at 1e-4 x RMS)")}
        log(f"{key}: {ms6:.3f} ms per pass, {out[key]['value']} Msps (plan {g6.plan}, warm {g6.warm})")
        del g6
    del d6, o6

    # iir_order4_bank: many streams of the same filter, each exact (one stream
    # per lane); 2^16 samples per stream per call, the streams' rows back to back
    bank_lines = {}
    for n_streams in (64, 256, 1024, 4096):
        nb_ = 1 << 16
        xs = np.resize(w6.input()[: 1 << 20], n_streams * nb_).astype(np.complex64)
        db = fd.DeviceBuffer.from_array(xs)
        ob = fd.DeviceBuffer(n_streams * nb_)
        bank = fd.IirBank(bq, aq, n_streams)

        def pass7():
            bank.apply_device(db, nb_, nb_, ob, nb_, stream)

        ms7 = timed_launches(fd, stream, pass7, 3)
        bank_lines[str(n_streams)] = {"value": round(n_streams * nb_ / (ms7 / 1e3) / 1e6, 1),
                                      "ms_per_call": round(ms7, 4)}
        log(f"iir_order4_bank: {n_streams} streams x 2^16: {ms7:.3f} ms, {bank_lines[str(n_streams)]['value']} Msps")
        del db, ob, bank, xs
    top = bank_lines["4096"]
    out["iir_order4_bank"] = {
        "value": top["value"], "unit": "output Msps", "streams": 4096, "by_streams": bank_lines,
        "roofline": {"bound": "hbm", "achieved": round(16 * top["value"] / 1e3, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(16 * top["value"] / 1e3 / HBM_PEAK_GBS, 4)},
        "note": "fdsp_iir_bank: 4th-order Butterworth on N independent streams, 2^16 cf32 samples each per call, "
                "every stream bit-identical to the reference's loop; value at 4096 streams, by_streams for 64..4096"}

    # fir_bank_256: 256 channels x 2^20, 512 real taps
    w4 = FirBank256()
    taps = w4.taps()
    fb = fd.FirBank(taps)
    d4 = fd.DeviceBuffer(w4.n_chan * w4.n)
    # each channel's own seeded samples (CWGN seed 3 at offset c * 2^20), made on the host cores
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(cpu_threads()) as pool:
        for c, xc in enumerate(pool.map(w4.input, range(w4.n_chan))):
            d4.upload(xc, offset=c * w4.n)
    o4 = fd.DeviceBuffer(w4.n_chan * w4.n)
    ms4 = timed_launches(fd, stream, lambda: fb.apply_device(d4, w4.n, w4.n, o4, w4.n, stream),
                         max(2, steps // 4))
    flops = 4.0 * w4.n_taps * w4.n * w4.n_chan
    tf = flops / (ms4 / 1e3) / 1e12
    out["fir_bank_256"] = {
        "value": round(tf, 2), "unit": "TFLOP/s", "ms_per_pass": round(ms4, 4),
        "roofline": {"bound": "mfma", "achieved": round(tf, 2), "peak": 157.3, "unit": "TFLOP/s",
                     "frac": round(tf / 157.3, 4)},
        "note": "4 flops per real tap per complex sample on FP32 matrix cores (Toeplitz form); each channel "
                "its own seeded CWGN samples"}
    log(f"fir_bank_256: {ms4:.3f} ms per pass, {tf:.1f} TFLOP/s")
    del d4, o4, fb
    return out


def bench_serving(fd):
    """Serving workload: Poisson arrivals (2000 req/s) of single-channel
    streams at 25 Msps, chunks U[4k, 256k], a quarter FM streams with a
    discriminator tail. Two runs: paced at the arrival times (latency at the
    offered load), then all requests at once (capacity)."""
    from libfalcon_dsp_amd.signals import cwgn
    from libfalcon_dsp_amd.workloads import Serving
    w = Serving()
    reqs = w.requests()
    pool = cwgn(w.seed, 1 << 23)

    def run(paced):
        srv = fd.Server(w.fs, max_streams=w.max_streams, max_chunk=w.chunk_hi, max_batch=1 << 22,
                        hbm_budget=w.hbm_budget)
        sids, tickets = {}, []
        t0 = time.perf_counter()
        for i, rq in enumerate(reqs):
            if paced:
                while True:
                    dt = t0 + rq.t_arrival - time.perf_counter()
                    if dt <= 0:
                        break
                    if dt > 2e-3:
                        time.sleep(dt - 1e-3)
            if rq.stream not in sids:
                sids[rq.stream] = srv.admit(w.stream_channel(rq.stream))
            off = (i * 2654435761) % (pool.size - rq.chunk)
            tickets.append((srv.submit(sids[rq.stream], pool[off:off + rq.chunk]), rq.chunk))
        res = [srv.result(t)[1] for t, _ in tickets]
        info = srv.info()
        srv.close()
        lat = [(r.done_ns - r.submit_ns) / 1e6 for r in res]
        span = (max(r.done_ns for r in res) - min(r.submit_ns for r in res)) / 1e9
        n_in = sum(m for _, m in tickets)
        n_out = sum(r.n_out for r in res)
        # where the time goes, per request (host steady clock): waiting for a
        # batch, packing and issuing it, the GPU's copy-in + kernels + copy-out
        # (until the worker sees it), delivery into the caller's buffer
        stages = {"queue": [(r.pack_ns - r.submit_ns) / 1e6 for r in res],
                  "pack_issue": [(r.issue_ns - r.pack_ns) / 1e6 for r in res],
                  "gpu": [(r.gpu_ns - r.issue_ns) / 1e6 for r in res],
                  "deliver": [(r.done_ns - r.gpu_ns) / 1e6 for r in res]}
        worst = sorted(range(len(lat)), key=lambda i: lat[i])[-max(1, len(lat) // 100):]
        return {"p50_ms": round(pct(lat, 0.5), 3), "p99_ms": round(pct(lat, 0.99), 3),
                "in_msps": round(n_in / span / 1e6, 1), "out_msps": round(n_out / span / 1e6, 2),
                "seconds": round(span, 3), "batches": info["batches"],
                "chunks_per_batch": round(info["chunks"] / max(1, info["batches"]), 2),
                "streams": len(sids),
                "stages_ms": {k: {"p50": round(pct(v, 0.5), 3), "p99": round(pct(v, 0.99), 3),
                                  "slowest_1pct_mean": round(sum(v[i] for i in worst) / len(worst), 3)}
                              for k, v in stages.items()}}

    paced = run(True)
    log(f"serving (paced): p50 {paced['p50_ms']} ms, p99 {paced['p99_ms']} ms, {paced['in_msps']} Msps in")
    log(f"serving (paced) stages: {paced['stages_ms']}")
    sat = run(False)
    log(f"serving (all at once): p50 {sat['p50_ms']} ms, p99 {sat['p99_ms']} ms, {sat['in_msps']} Msps in")
    return {"workload": f"{len(reqs)} requests, Poisson {w.rate_hz:g}/s over {w.max_streams} streams at "
                        f"{w.fs / 1e6:g} Msps, chunks U[{w.chunk_lo}, {w.chunk_hi}] from host memory, "
                        f"{w.fm_share:g} of streams FM with a discriminator tail, budget {w.hbm_budget >> 30} GiB",
            "latency": "submit to result in the caller's host buffer (server clock)",
            "paced": paced, "saturated": sat}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the per-channel plain GPU form")
    ap.add_argument("--latency-chunk", type=int, default=1 << 18)
    ap.add_argument("--latency-chunks", type=int, default=200)
    ap.add_argument("--seg", type=int, default=0,
                    help="fused-kernel work split (1 strided, or a segment length; 0: autotune)")
    ap.add_argument("--block", type=int, default=0, help="fused-kernel block threads (0: autotune)")
    ap.add_argument("--no-others", action="store_true", help="skip the resampler, NCO and FIR-bank workloads")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving workload")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 3 device-to-device copies of 2^24 complex samples (128 MiB), the known-size "
                         "copy that scales the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()

    if args.copy_calibration:
        import libfalcon_dsp_amd as fd
        n = 1 << 24
        a, b = fd.DeviceBuffer(n), fd.DeviceBuffer(n)
        a.zero()
        st = fd.Stream()
        for _ in range(3):
            fd.lib().fdsp_memcpy_d2d(b.ptr, a.ptr, n * 8, st.h)
        st.sync()
        print(json.dumps({"copy_bytes_per_dispatch": n * 8, "dispatches": 3}), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the channelizer shards by "
              "channel across GPUs with one broadcast of the input, which this build does not run.",
              file=sys.stderr)
        sys.exit(2)

    import libfalcon_dsp_amd as fd
    from libfalcon_dsp_amd.workloads import Chan32

    w = Chan32()
    log("generating the chan_32 workload (2^24 samples, 32 channels)")
    chans = w.channels()
    x = w.input()
    n = x.size

    g = fd.MultiRateChannelizer(w.fs, chans, batched=True)  # FP32 (the default arithmetic)
    kinfo = g.kernel_info()
    stream = fd.Stream()
    d_in = fd.DeviceBuffer.from_array(x)
    trials = None
    if args.seg or args.block:
        s0, b0 = g.get_tiling()
        g.set_tiling(args.seg or s0, args.block or b0)
    else:
        trials = g.autotune(d_in, n, reps=5)
        log(f"{kinfo['kernel']}: autotune: " + ", ".join(
            f"{t['segment_len']}/{t['block_threads']}{'m' if t['measured_weights'] else ''}"
            f"{'f' if t['fused'] else ''}: {t['us']} us" for t in trials))
    seg, block = g.get_tiling()
    outs = [fd.DeviceBuffer(m) for m in g.max_out(n)]
    g.graph_create(n, d_in, outs)
    fd.device_sync()

    log(f"warmup {args.warmup} graph launches (segment {seg}, block {block})")
    for _ in range(args.warmup):
        g.graph_launch(stream)
    stream.sync()

    # throughput: K back-to-back launches
    e0, e1 = fd.Event(), fd.Event()
    total_out = 0
    e0.record(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.graph_launch(stream)
        total_out += sum(g.graph_counts())
    e1.record(stream)
    e1.sync()
    t1 = time.perf_counter()
    ms = e0.elapsed_ms(e1)
    ms_per_step = ms / args.steps
    value = total_out / (ms / 1e3) / 1e6
    log(f"throughput: {ms_per_step:.4f} ms/step, {value:.1f} Msps aggregate output "
        f"(host clock {1e3 * (t1 - t0) / args.steps:.4f} ms/step)")

    # kernel time of the fused kernel: the eager form launches exactly that one
    # kernel per call, so events on its stream bracket it
    # (back-to-back launches between one pair of events on the kernel's
    # stream: the host queues the next launch while the current one runs, so
    # the average is kernel time, not host latency)
    out_per = []
    k0, k1 = fd.Event(), fd.Event()
    g.apply_device(d_in, n, outs, stream)  # first call uploads the output table
    stream.sync()
    k0.record(stream)
    for _ in range(args.steps):
        out_per.append(sum(g.apply_device(d_in, n, outs, stream)))
    k1.record(stream)
    k1.sync()
    k_avg = k0.elapsed_ms(k1) / args.steps
    alg_bytes = 8 * n + 8 * statistics.mean(out_per)
    achieved = alg_bytes / (k_avg / 1e3) / 1e9
    log(f"{kinfo['kernel']}: avg {k_avg * 1e3:.1f} us, algorithmic {alg_bytes / 1e6:.1f} MB -> {achieved:.0f} GB/s")
    traffic, traffic_src = measured_traffic(kinfo["kernel"])

    # p50/p99 chunk latency: streaming chunks of --latency-chunk through a graph
    lc = args.latency_chunk
    gl = fd.MultiRateChannelizer(w.fs, chans, batched=True)
    gl.set_tiling(seg, block)
    lin = fd.DeviceBuffer(lc)
    louts = [fd.DeviceBuffer(m) for m in gl.max_out(lc)]
    gl.graph_create(lc, lin, louts)
    lat = []
    a0, a1 = fd.Event(), fd.Event()
    for k in range(args.latency_chunks + 5):
        off = (k * lc) % (n - lc)
        fd.lib().fdsp_memcpy_d2d(lin.ptr, d_in.at(off), lc * 8, stream.h)
        a0.record(stream)
        gl.graph_launch(stream)
        a1.record(stream)
        a1.sync()
        if k >= 5:
            lat.append(a0.elapsed_ms(a1))
    lat.sort()
    p50 = lat[len(lat) // 2]
    p99 = lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))]
    log(f"chunk latency ({lc} samples, device-resident): p50 {p50 * 1e3:.1f} us, p99 {p99 * 1e3:.1f} us")

    plain = None
    if not args.no_plain:
        gp = fd.MultiRateChannelizer(w.fs, chans, batched=False)
        pouts = [fd.DeviceBuffer(m) for m in gp.max_out(n)]
        gp.apply_device(d_in, n, pouts, stream)
        stream.sync()
        psteps = max(2, min(5, args.steps))
        p0, p1 = fd.Event(), fd.Event()
        tot = 0
        p0.record(stream)
        for _ in range(psteps):
            tot += sum(gp.apply_device(d_in, n, pouts, stream))
        p1.record(stream)
        p1.sync()
        pms = p0.elapsed_ms(p1)
        plain = {"value": round(tot / (pms / 1e3) / 1e6, 1), "unit": "Msps",
                 "ms_per_step": round(pms / psteps, 4),
                 "form": "per-channel: one multi-shift NCO launch writing 32 full-rate copies, "
                         "then 32 resampler launches"}
        log(f"plain per-channel GPU form: {plain['value']} Msps ({plain['ms_per_step']} ms/step)")
        del gp, pouts

    others = None if args.no_others else bench_other_configs(fd, stream, args.steps)
    if not args.no_others:
        # the same workload on the channel-group kernel in FP32, and on the opt-in
        # reduced-precision form (bf16 hi/lo planes), each labelled by its arithmetic
        gf = chan32_graph_rate(fd, w, chans, d_in, stream, args.steps, args.warmup, "f32", kernel="groups")
        gf.update({"dtype": "f32", "note": "channel-group kernel (FDSP_KERNEL_GROUPS_F32): each group family "
                                           "stages its own tiles; FP32 MFMA and VALU"})
        log(f"chan_32 on {gf['kernel']} (f32): {gf['value']} Msps, kernel {gf['kernel_avg_us']} us")
        gb = chan32_graph_rate(fd, w, chans, d_in, stream, args.steps, args.warmup, "bf16x2")
        gb.update({"dtype": "bf16x2-split", "note": "FDSP_CHAN_BATCHED_BF16X2, opt-in: groups of 32-tap windows "
                                                    "with an even down below 32 multiply bf16 hi/lo splits (about "
                                                    "16 significant bits, lo*lo dropped, FP32 accumulation); not "
                                                    "an FP32 result"})
        log(f"chan_32 bf16x2: {gb['value']} Msps, kernel {gb['kernel_avg_us']} us")
        others["chan_32_groups_f32"] = gf
        others["chan_32_bf16x2"] = gb
    serving = None if args.no_serving else bench_serving(fd)

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (reference channelizer on the host cores)")
        cpu = cpu_baseline(w.fs, chans, x, 1 << 22, n)
        log(f"CPU baseline: {cpu['value']} Msps on {cpu['cores']} threads, "
            f"{cpu['single_thread_value']} Msps single thread ({cpu['kind']})")
        if others and "iir_order4" in others:
            others["iir_order4"]["cpu_baseline"] = iir_cpu_baseline()
            log(f"IIR CPU baseline: {others['iir_order4']['cpu_baseline']['value']} Msps single thread")
            # the bank's streams are independent: the host runs one per thread
            if "iir_order4_bank" in others:
                one = others["iir_order4"]["cpu_baseline"]
                th = cpu_threads()
                others["iir_order4_bank"]["cpu_baseline"] = dict(
                    one, value=round(one["value"] * th, 2), cores=th,
                    sample=one["sample"] + f", scaled by {th} threads (independent streams)")
        if others:
            for key, v in other_cpu_baselines().items():
                if key in others:
                    others[key]["cpu_baseline"] = v
                    log(f"{key} CPU baseline: {v['value']} {v['unit']} on {v['cores']} threads ({v['kind']})")

    rec = {
        "metric": "aggregate output Msps (32-ch multi-rate channelizer) + p99 chunk latency",
        "value": round(value, 1),
        "unit": "Msps",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded CWGN + 32 tones, splitmix64/Box-Muller)",
        "config": {"workload": "chan_32: 2^24 cf32 samples at 20 Msps, 32 channels, outputs "
                               "{12.5k,25k,100k,250k,1M,2.5M} sps, 32 Kaiser taps, one fused FP32 launch per "
                               "step in a HIP graph",
                   "input_samples": n, "channels": len(chans), "outputs_per_step": int(round(total_out / args.steps)),
                   "kernel": kinfo["kernel"], "work_groups": kinfo["groups"], "mfma_groups": kinfo["mfma_groups"],
                   "segment_len": seg, "block_threads": block, "autotune_trials": trials},
        "p50_chunk_latency_ms": round(p50, 5),
        "p99_chunk_latency_ms": round(p99, 5),
        "latency_chunk_samples": lc,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": kinfo["kernel"], "kernel_avg_us": round(k_avg * 1e3, 2),
                     "algorithmic_bytes": int(alg_bytes)},
        "plain_gpu": plain,
        "batched_vs_plain": round(value / plain["value"], 1) if plain else None,
        "cpu_baseline": cpu,
        "serving": serving,
        "other_workloads": others,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
