"""Writes tests/golden/fm_taps.npz: the two resampling filters the reference's
FM demodulator reads from its predefined table (get_resample_fir_params with
FILTER_TAPS_64 and REMEZ, reference src/transform/falcon_dsp_fir_filter.cc:111-130,
as used by src/signals/falcon_dsp_fm_demod.cc:125-186): 1,140,000 -> 228,000
sps and 228,000 -> 45,600 sps, each (up, down, taps). The values come from
oracle/_ref, the reference's CPU sources compiled in this container; the
GPU tests read this file instead of the reference.

    python tests/golden/make_fm_taps.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

RATES = {"iso": (1_140_000, 228_000), "mono": (228_000, 45_600)}


def main():
    out = {}
    for name, (fin, fout) in RATES.items():
        up, down, taps = O.ref_resample_fir_params(fin, fout)
        out[f"{name}_rates"] = np.array([fin, fout, up, down], dtype=np.int64)
        out[f"{name}_taps"] = taps.astype(np.complex64)
        print(name, fin, "->", fout, "up", up, "down", down, "taps", taps.size)
    np.savez(os.path.join(HERE, "fm_taps.npz"), **out)


if __name__ == "__main__":
    main()
