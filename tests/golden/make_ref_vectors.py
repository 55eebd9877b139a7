"""Write ref_vectors.npz: outputs of the reference's own CPU classes
(oracle/_ref/libfdsp_ref.so, compiled from the reference sources by
oracle/Makefile) on small seeded inputs, streamed in uneven chunks.

    python tests/golden/make_ref_vectors.py

The inputs are regenerated from the seeds stored beside the outputs
(libfalcon_dsp_amd.signals), so the fixture holds only parameters, seeds and
expected outputs. The tests compare the oracle restatement (and the GPU path)
against these vectors on machines where the reference is absent.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from libfalcon_dsp_amd.signals import cwgn, kaiser_sinc  # noqa: E402
from oracle import oracle as O  # noqa: E402

N = 6000
CHUNKS = [2047, 1, 3, 1500, 2449]


def chunked(obj, x):
    out, i, k = [], 0, 0
    while i < x.size:
        m = min(CHUNKS[k % len(CHUNKS)], x.size - i)
        out.append(obj.apply(x[i:i + m]))
        i += m
        k += 1
    return out


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libfdsp_ref.so is not built: run make -C oracle (needs the reference sources)")
    rec = {}
    # frequency shift (reference src/transform/falcon_dsp_freq_shift.cc)
    fs_cases = [(1_000_000, 1234), (1_000_000, -250_000), (25_000_000, 3_333_333), (48_000, 7)]
    for i, (fs, f) in enumerate(fs_cases):
        x = cwgn(100 + i, N)
        y = np.concatenate(chunked(O.RefFreqShift(fs, f), x))
        rec[f"fs{i}_params"] = np.array([fs, f, 100 + i], np.int64)
        rec[f"fs{i}_y"] = y
    # polyphase resampler (src/transform/falcon_dsp_polyphase_resampler.cc)
    rs_cases = [(1, 2, np.array([0.5, 1.0, 2.0])), (4, 5, kaiser_sinc(96, 0.1, 8.0, gain=4)),
                (3, 7, kaiser_sinc(50, 0.5 / 7, 6.0, gain=3)), (2, 1, kaiser_sinc(20, 0.25, 5.0, gain=2)),
                (1, 8, kaiser_sinc(33, 0.0625, 8.0)), (5, 3, kaiser_sinc(61, 0.1, 7.0, gain=5))]
    for i, (up, down, h) in enumerate(rs_cases):
        x = cwgn(200 + i, N)
        taps = np.asarray(h, np.complex64)
        r = O.RefResampler(up, down, taps)
        ys = chunked(r, x)
        rec[f"rs{i}_params"] = np.array([up, down, 200 + i], np.int64)
        rec[f"rs{i}_taps"] = taps
        rec[f"rs{i}_counts"] = np.array([y.size for y in ys], np.int64)
        rec[f"rs{i}_y"] = np.concatenate(ys)
    # FIR with complex taps (src/transform/falcon_dsp_fir_filter.cc)
    for i, m in enumerate((3, 17, 64)):
        u = cwgn(300 + i, m)
        taps = (u / np.sqrt(m)).astype(np.complex64)
        x = cwgn(310 + i, N)
        rec[f"fir{i}_taps"] = taps
        rec[f"fir{i}_seed"] = np.array([310 + i], np.int64)
        rec[f"fir{i}_y"] = np.concatenate(chunked(O.RefFir(taps), x))
    # polar discriminator (src/transform/falcon_dsp_polar_discriminator.cc)
    x = cwgn(400, N)
    rec["disc_seed"] = np.array([400], np.int64)
    rec["disc_y"] = np.concatenate(chunked(O.RefDiscriminator(), x))
    # multi-rate channelizer (src/transform/falcon_dsp_multi_rate_channelizer.cc):
    # 2 Msps, five channels as in the reference's _016 recipe, plus one up > 1
    fs = 2_000_000
    chans = [(250_000, 60_000), (100_000, -350_000), (500_000, 200_000), (40_000, -20_001), (300_000, 123_457)]
    descs = []
    for rate, shift in chans:
        g = int(np.gcd(fs, rate))
        up, down = rate // g, fs // g
        taps = kaiser_sinc(21 * max(up, 1), 0.5 / max(up, down), 8.0, gain=up).astype(np.complex64)
        descs.append(dict(output_sample_rate_in_sps=rate, freq_shift_in_hz=shift, up_rate=up, down_rate=down,
                          taps=taps))
    x = cwgn(500, 4 * N)
    ch = O.RefChannelizer(fs, descs)
    per = [[] for _ in descs]
    i, k = 0, 0
    while i < x.size:
        m = min(CHUNKS[k % len(CHUNKS)], x.size - i)
        for c, y in enumerate(ch.apply(x[i:i + m])):
            per[c].append(y)
        i += m
        k += 1
    rec["chan_fs"] = np.array([fs, 500], np.int64)
    for c, d in enumerate(descs):
        rec[f"chan{c}_params"] = np.array([d["output_sample_rate_in_sps"], d["freq_shift_in_hz"], d["up_rate"],
                                           d["down_rate"]], np.int64)
        rec[f"chan{c}_taps"] = d["taps"]
        rec[f"chan{c}_counts"] = np.array([y.size for y in per[c]], np.int64)
        rec[f"chan{c}_y"] = np.concatenate(per[c])
    out = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(out, chunks=np.array(CHUNKS, np.int64), n=np.array([N], np.int64), **rec)
    print(f"wrote {out} ({os.path.getsize(out) / 1024:.0f} KiB, {len(rec)} arrays)")


if __name__ == "__main__":
    main()
