"""Complex int16 path on the CPU: the oracle's ci16 overloads and the drop-in's
sample-file functions.

The reference reads complex int16 samples as floats, runs the float path and
builds its int16 outputs from float parts: truncation toward zero
(src/transform/falcon_dsp_freq_shift.cc:133-162, falcon_dsp_fir_filter.cc:151-175,
falcon_dsp_polar_discriminator.cc:117-132). Pinned here by:
- the compiled reference classes (oracle/_ref) on the same chunked inputs: the
  shift and FIR outputs equal sample for sample except where a float lands on
  a truncation boundary (at most 1 LSB, in under 0.1% of samples);
- exact cases: identity and half-gain taps (no rounding), the truncation table
  of orc_i16, and saturation outside the int16 range (a build choice: the
  reference's conversion is undefined there).
Sample files (falcon_dsp_utils.cc:301-811): the drop-in writer's bytes equal
the reference writer's (binary and ASCII); the drop-in reader reads what the
reference wrote; and the reference's int16 binary reader corrupts values
whose low byte has its top bit set (.cc:449,696-697), which the drop-in's does not.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from libfalcon_dsp_amd.signals import cwgn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FILES_BIN = os.path.join(ROOT, "tests", "dropin", "test_files")
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")


def ci16_noise(seed, n, scale):
    x = cwgn(seed, n)
    return np.stack([np.round(x.real * scale), np.round(x.imag * scale)], axis=1).clip(-32768, 32767).astype(np.int16)


def near_int(a, b, what):
    d = np.abs(a.astype(np.int32) - b.astype(np.int32))
    assert d.max(initial=0) <= 1, f"{what}: max diff {d.max()}"
    assert np.count_nonzero(d) <= max(1, a.size // 1000), f"{what}: {np.count_nonzero(d)} of {a.size} differ"


def chunked(obj, x, sizes):
    out, i = [], 0
    for s in sizes:
        out.append(obj.apply_ci16(x[i:i + s]))
        i += s
    return np.concatenate(out)


def test_orc_i16_truncates_toward_zero_and_saturates():
    L = O.lib()
    table = {0.0: 0, 0.99: 0, -0.99: 0, 1.5: 1, -1.5: -1, 32766.9: 32766, -32767.9: -32767, 32767.0: 32767,
             40000.0: 32767, -32768.0: -32768, -1e9: -32768, float("nan"): -32768}
    for v, want in table.items():
        assert L.orc_i16(v) == want, v


def test_oracle_ci16_exact_cases():
    x = np.array([[3, -3], [5, -5], [-1, 1], [32767, -32768], [0, 7]], dtype=np.int16)
    half = O.Fir(np.array([0.5], dtype=np.complex64)).apply_ci16(x)
    np.testing.assert_array_equal(half, [[1, -1], [2, -2], [0, 0], [16383, -16384], [0, 3]])
    ident = O.Fir(np.array([1.0], dtype=np.complex64)).apply_ci16(x)
    np.testing.assert_array_equal(ident, x)
    gain4 = O.Fir(np.array([4.0], dtype=np.complex64)).apply_ci16(np.array([[20000, -20000]], dtype=np.int16))
    np.testing.assert_array_equal(gain4, [[32767, -32768]])
    # the discriminator of int16 input is the float discriminator of the same values
    y = ci16_noise(1, 500, 3000)
    np.testing.assert_array_equal(O.Discriminator().apply_ci16(y),
                                  O.Discriminator().apply(y[:, 0].astype(np.float32) + 1j * y[:, 1].astype(np.float32)))
    # the shift of an int16 chunk is the truncated float shift, one counter step per sample
    s16, sf = O.FreqShift(2_000_000, 123_457), O.FreqShift(2_000_000, 123_457)
    yf = sf.apply(y[:, 0].astype(np.float32) + 1j * y[:, 1].astype(np.float32))
    want = np.stack([np.trunc(yf.real), np.trunc(yf.imag)], axis=1).astype(np.int16)
    np.testing.assert_array_equal(s16.apply_ci16(y), want)
    assert s16.samples_handled == sf.samples_handled


@needs_ref
@pytest.mark.parametrize("seed", range(3))
def test_oracle_ci16_vs_reference_live(seed):
    u = np.random.default_rng(seed)
    x = ci16_noise(seed + 10, 20_000, float(u.uniform(100, 9000)))
    sizes = [int(v) for v in u.integers(1, 6000, 8)]
    sizes.append(x.shape[0] - sum(sizes) if sum(sizes) < x.shape[0] else 0)
    sizes = [s for s in sizes if s > 0]
    x = x[:sum(sizes)]
    fs, f = 20_000_000, int(u.integers(-9_000_000, 9_000_000)) or 1
    near_int(chunked(O.FreqShift(fs, f), x, sizes), chunked(O.RefFreqShift(fs, f), x, sizes), "freq shift")
    # taps normalised so every output stays inside int16 (outside it the
    # reference's conversion is undefined)
    m = int(u.integers(1, 40))
    h = ((u.standard_normal(m) + 1j * u.standard_normal(m)) / np.sqrt(2 * m)).astype(np.complex64)
    xs = ci16_noise(seed + 20, x.shape[0], 500.0)
    assert np.abs(xs.astype(np.float64)).max() * np.abs(h).sum() * np.sqrt(2) < 32767
    near_int(chunked(O.Fir(h), xs, sizes), chunked(O.RefFir(h), xs, sizes), "fir")
    a, b = chunked(O.Discriminator(), x, sizes), chunked(O.RefDiscriminator(), x, sizes)
    assert a.size == b.size == x.shape[0] - 1
    assert np.max(np.abs(np.angle(np.exp(1j * (a.astype(np.float64) - b))))) < 1e-5


# --------------------------------------------------------------------------
# sample files
# --------------------------------------------------------------------------
KINDS = {"ci16": (0, np.int16, 2), "cf32": (1, np.float32, 2), "i16": (2, np.int16, 1), "f32": (3, np.float32, 1)}


def _data(kind):
    u = np.random.default_rng(5)
    _, dt, parts = KINDS[kind]
    n = 300 * parts
    if dt == np.int16:
        v = u.integers(-32768, 32768, n).astype(np.int16)
        v[:6] = [384, -384, 0x7F80, -32768, 32767, 128]  # low bytes with the top bit set
    else:
        v = (u.standard_normal(n) * 1e3).astype(np.float32)
    return v


def _convert(kind, fin, pin, fout, pout):
    return subprocess.run([FILES_BIN, kind, fin, pin, fout, pout], capture_output=True).returncode


@needs_ref
@pytest.mark.parametrize("kind", list(KINDS))
def test_sample_files_match_the_reference(kind, tmp_path):
    assert os.path.exists(FILES_BIN), "build first: python -m libfalcon_dsp_amd.build"
    code, dt, parts = KINDS[kind]
    v = _data(kind)
    n_items = v.size // parts
    truth = tmp_path / "truth.bin"
    v.astype(np.dtype(dt).newbyteorder("<")).tofile(truth)
    L = O.ref()
    for ascii, ext in ((0, "bin"), (1, "txt")):
        ref_path = tmp_path / f"ref.{ext}"
        assert L.ref_write_file(code, ascii, str(ref_path).encode(), v.ctypes.data, n_items) == 0
        ours = tmp_path / f"ours.{ext}"
        fmt = "ascii" if ascii else "bin"
        if dt == np.float32:
            # binary floats cannot be read back (false in both); write from the text form
            assert _convert(kind, "bin", str(truth), fmt, str(ours)) == (2 if parts == 1 else 0)
            if parts == 1:
                assert L.ref_read_file(code, 0, str(truth).encode(), None, 0) == -1
                continue
        else:
            assert _convert(kind, "bin", str(truth), fmt, str(ours)) == 0
        assert ours.read_bytes() == ref_path.read_bytes(), f"{kind} {fmt}: writer bytes differ"
    if dt == np.float32 and parts == 1:
        # ASCII floats: the reference's text, read by the drop-in, written back as text, is the same text
        ref_txt = tmp_path / "ref.txt"
        back = tmp_path / "back.txt"
        assert _convert(kind, "ascii", str(ref_txt), "ascii", str(back)) == 0
        assert back.read_bytes() == ref_txt.read_bytes()
        return
    # the drop-in reads the reference's binary file back to the exact values
    back = tmp_path / "back.bin"
    assert _convert(kind, "bin", str(tmp_path / "ref.bin"), "bin", str(back)) == 0
    assert back.read_bytes() == truth.read_bytes()
    # and the reference's own reader on the same file
    got = np.empty(v.size, dtype=dt)
    assert L.ref_read_file(code, 0, str(truth).encode(), got.ctypes.data, n_items) == n_items
    if dt == np.int16:
        bad = (v.view(np.uint16) & 0x80) != 0
        assert np.array_equal(got[~bad], v[~bad])
        assert not np.array_equal(got[bad], v[bad]), "the reference's int16 reader no longer sign-extends"
        assert got[0] == -128  # 384 = 0x0180 reads as 0xFF80
    else:
        assert np.array_equal(got, v)
