"""General IIR filter on the CPU: the oracle pinned before the GPU is checked against it.

Pinned by:
- the reference's own known-answer tests (test/tests/transform/
  falcon_dsp_iir_filter_cpp_unit_tests.cc:147-260): an unstable first-order
  filter with a[0] = 0.5 on int16 and float samples, a second-order feedback
  and a feedforward-only filter;
- the compiled reference (oracle/_ref): falcon_dsp_iir_filter on all four
  sample types, chunked, with real and complex a[0] (which the reference
  divides through with std::complex<float> division), equal bit for bit;
- the division itself: the reference's 1/a0 read back from an impulse
  response equals Smith's method in double and not the naive float formula;
- scipy.signal.lfilter in FP64 (tests/test_fm_cpu.py).
"""
import numpy as np
import pytest

from libfalcon_dsp_amd.signals import cwgn
from oracle import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")


def chunked(f, x, sizes):
    out, i = [], 0
    for s in sizes:
        out.append(f(x[i:i + s]))
        i += s
    return np.concatenate(out)


def test_reference_known_answers():
    # cpp_basic_filter_uint16_t_000: b = [1], a = [0.5, 1] -> b = [2], a = [., 2]
    f = O.Iir([1.0], [0.5, 1.0])
    got = f.apply_ci16(np.array([[0, 0], [1, 0], [1, 0], [1, 0], [1, 0]], np.int16))
    np.testing.assert_array_equal(got, [[0, 0], [2, 0], [-2, 0], [6, 0], [-10, 0]])
    # cpp_basic_filter_float_000
    got = O.Iir([1.0], [0.5, 1.0]).apply(np.array([0.25, 1, 1, 1, 1], np.complex64))
    np.testing.assert_allclose(got, [0.5, 1.0, 0.0, 2.0, -2.0], atol=1e-4)
    # cpp_basic_filter_float_001
    got = O.Iir([1.0], [1.0, 0.9, 0.8]).apply(np.array([0, 1, 1, 1, 1], np.complex64))
    np.testing.assert_allclose(got, [0.0, 1.0, 0.1, 0.11, 0.821], atol=1e-4)
    # cpp_basic_filter_float_002 (feedforward only)
    got = O.Iir([0.3, 0.16, 0.21], [1.0]).apply(np.array([0, 10, 15, 11, 13], np.complex64))
    np.testing.assert_allclose(got, [0.0, 3.0, 6.1, 7.8, 8.81], atol=1e-4)


def random_filter(u, nb, nfb, complex_coeffs, a0):
    b = u.standard_normal(nb) * 0.5
    if complex_coeffs:
        b = b + 1j * u.standard_normal(nb) * 0.5
    # poles inside radius 0.9, expanded to a[1:]
    r = u.uniform(0.2, 0.9, nfb)
    ang = u.uniform(-np.pi, np.pi, nfb)
    poles = r * np.exp(1j * ang)
    if not complex_coeffs:
        poles = np.concatenate([poles[: nfb // 2], np.conj(poles[: nfb // 2])])
        if nfb % 2:
            poles = np.append(poles, r[-1])
    a = np.poly(poles) if nfb else np.array([1.0])
    a = (a if complex_coeffs else a.real) * a0
    return np.asarray(b, np.complex64), np.asarray(a, np.complex64)


@needs_ref
@pytest.mark.parametrize("seed", range(6))
def test_oracle_equals_reference_all_types(seed):
    u = np.random.default_rng(seed)
    nb, nfb = int(u.integers(1, 9)), int(u.integers(0, 9))
    cplx = bool(seed % 2)
    a0 = [1.0, 2.0, 0.7, 1.3 - 0.4j, 0.5 + 2.0j, -1.5][seed]
    b, a = random_filter(u, nb, nfb, cplx, a0)
    sizes = [1, 2, 9, 500, 3000, 6488]
    n = sum(sizes)
    x = cwgn(seed, n)
    np.testing.assert_array_equal(chunked(O.Iir(b, a).apply, x, sizes), chunked(O.RefIir(b, a).apply, x, sizes))
    xr = x.real.copy()
    np.testing.assert_array_equal(chunked(O.Iir(b, a).apply_f32, xr, sizes),
                                  chunked(O.RefIir(b, a).apply_f32, xr, sizes))
    xi = np.round(np.stack([x.real, x.imag], 1) * 300).astype(np.int16)
    np.testing.assert_array_equal(chunked(O.Iir(b, a).apply_ci16, xi, sizes),
                                  chunked(O.RefIir(b, a).apply_ci16, xi, sizes))
    np.testing.assert_array_equal(chunked(O.Iir(b, a).apply_i16, xi[:, 0].copy(), sizes),
                                  chunked(O.RefIir(b, a).apply_i16, xi[:, 0].copy(), sizes))
    # one object, types mixed call to call, state shared
    f, r = O.Iir(b, a), O.RefIir(b, a)
    for part in (slice(0, 700), slice(700, 1500)):
        np.testing.assert_array_equal(f.apply(x[part]), r.apply(x[part]))
        np.testing.assert_array_equal(f.apply_f32(xr[part]), r.apply_f32(xr[part]))
    f.reset()
    r.reset()
    np.testing.assert_array_equal(f.apply(x[:999]), r.apply(x[:999]))


@needs_ref
def test_a0_division_is_smith_in_double():
    a0 = np.complex64(0.5312527418136597 - 1.052757978439331j)
    x = np.zeros(3, np.complex64)
    x[0] = 1
    y0 = O.RefIir([1.0], [a0, 0.3]).apply(x)[0]
    c, d = float(a0.real), float(a0.imag)
    r = d / c
    den = d * r + c
    smith = complex(np.float32((0 * r + 1) / den), np.float32((0 - 1 * r) / den))
    assert y0 == np.complex64(smith)
    assert O.Iir([1.0], [a0, 0.3]).apply(x)[0] == y0
    # the naive float formula rounds differently for this divisor
    f = np.float32
    den_f = f(f(a0.real * a0.real) + f(a0.imag * a0.imag))
    naive = complex(f(a0.real / den_f), f(-a0.imag / den_f))
    assert np.complex64(naive) != y0
