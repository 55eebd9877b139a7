"""Python mirror of the reference's operator classes, over the C ABI.

Each class keeps the reference's name, argument meaning and error behaviour
(methods return what the reference returns: bools, counts, arrays) while the
work runs in the gfx950 kernels of libfdsp_hip.so:

=============================  ======================================================
this module                    reference (hdr/transform/*.h)
=============================  ======================================================
FreqShift                      falcon_dsp_freq_shift{,_cuda}      (freq_shift.h:98-121)
PolyphaseResampler             falcon_dsp_polyphase_resampler{,_cuda} (.h:168-195)
FirFilter                      falcon_dsp_fir_filter{,_cuda}      (fir_filter.h:119-140)
FirBank                        multi-channel falcon_dsp_fir_filter with real taps
PolarDiscriminator             falcon_dsp_polar_discriminator     (.h:97-114)
MultiRateChannelizer           falcon_dsp_multi_rate_channelizer{,_cuda} (.h:97-134)
freq_shift / resample /        free functions (freq_shift.h:85, resample.h:85-95,
upfirdn / fir_filter /         fir_filter.h:87-91, polar_discriminator.h:81-85)
polar_discriminator
=============================  ======================================================

Host arrays are numpy complex64 (interleaved float pairs, the
std::complex<float> layout). ``DeviceBuffer`` holds device memory for the
asynchronous ``*_device`` forms.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import _native as N
from ._native import check, lib

_vp = C.c_void_p

# MultiRateChannelizer.set_tiling work splits other than a segment length
# (FDSP_TILING_CLAIM / FDSP_TILING_STRIDED in include/fdsp.h)
TILING_CLAIM = 0
TILING_STRIDED = 1


def _cf32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.complex64))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_vp)


# ---------------------------------------------------------------------------
# device plumbing
# ---------------------------------------------------------------------------
def device_count() -> int:
    n = C.c_int(0)
    check(lib().fdsp_device_count(C.byref(n)))
    return n.value


class Stream:
    def __init__(self):
        self.h = _vp()
        check(lib().fdsp_stream_create(C.byref(self.h)))

    def sync(self):
        check(lib().fdsp_stream_sync(self.h))

    def close(self):
        if self.h:
            lib().fdsp_stream_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.h = _vp()
        check(lib().fdsp_event_create(C.byref(self.h)))

    def record(self, stream: Optional[Stream] = None):
        check(lib().fdsp_event_record(self.h, stream.h if stream else None))

    def sync(self):
        check(lib().fdsp_event_sync(self.h))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib().fdsp_event_elapsed_ms(self.h, end.h, C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            if self.h:
                lib().fdsp_event_destroy(self.h)
        except Exception:
            pass


class DeviceBuffer:
    """Device allocation of ``count`` elements of ``dtype``."""

    def __init__(self, count: int, dtype=np.complex64):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        self.ptr = _vp()
        check(lib().fdsp_malloc(C.byref(self.ptr), max(self.nbytes, 1)))

    @classmethod
    def from_array(cls, a: np.ndarray, stream: Optional[Stream] = None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.size, a.dtype)
        b.upload(a, stream)
        return b

    def upload(self, a: np.ndarray, stream: Optional[Stream] = None, offset: int = 0):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if offset + a.size > self.count:
            raise ValueError("upload past the end of the buffer")
        dst = _vp(self.ptr.value + offset * self.dtype.itemsize)
        check(lib().fdsp_memcpy_h2d(dst, _ptr(a), a.nbytes, stream.h if stream else None))

    def download(self, count: Optional[int] = None, offset: int = 0, stream: Optional[Stream] = None) -> np.ndarray:
        count = self.count - offset if count is None else int(count)
        out = np.empty(count, dtype=self.dtype)
        if count:
            src = _vp(self.ptr.value + offset * self.dtype.itemsize)
            check(lib().fdsp_memcpy_d2h(_ptr(out), src, out.nbytes, stream.h if stream else None))
            if stream:
                stream.sync()
        return out

    def at(self, offset: int) -> _vp:
        return _vp(self.ptr.value + offset * self.dtype.itemsize)

    def zero(self, stream: Optional[Stream] = None):
        check(lib().fdsp_memset(self.ptr, 0, self.nbytes, stream.h if stream else None))

    def free(self):
        if self.ptr:
            lib().fdsp_free(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_sync():
    check(lib().fdsp_device_sync())


# complex int16 samples: an (n, 2) int16 array of (re, im) pairs, the layout
# of std::complex<int16_t> and fdsp_ci16
CI16 = np.dtype([("re", np.int16), ("im", np.int16)])


def _ci16(x) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.int16))
    if a.ndim != 2 or a.shape[1] != 2:
        raise ValueError("complex int16 samples are an (n, 2) int16 array of (re, im)")
    return a


# ---------------------------------------------------------------------------
# rate helpers
# ---------------------------------------------------------------------------
def calculate_gcd(a: int, b: int) -> int:
    """calculate_gcd (reference src/utilities/falcon_dsp_utils.cc:123-126)."""
    return int(lib().fdsp_gcd(a, b))


def calculate_lcm(a: int, b: int) -> int:
    """calculate_lcm (reference src/utilities/falcon_dsp_utils.cc:133-137)."""
    return int(lib().fdsp_lcm(a, b))


def rat_approx(f: float, max_denom: int):
    """rat_approx (reference src/utilities/falcon_dsp_utils.cc:231-294) -> (num, denom)."""
    n, d = C.c_int64(), C.c_int64()
    lib().fdsp_rat_approx(float(f), int(max_denom), C.byref(n), C.byref(d))
    return n.value, d.value


def get_freq_shift_params(fs: int, f_hz: int):
    """falcon_dsp_freq_shift::get_freq_shift_params (reference freq_shift.cc:194-235)
    -> (rollover index, angular frequency in rad/sample)."""
    r, w = C.c_uint32(), C.c_double()
    check(lib().fdsp_freq_shift_params(fs, f_hz, C.byref(r), C.byref(w)))
    return r.value, w.value


def reduce_rates(in_rate: int, out_rate: int):
    """Exact up/down for a rational resample: (out/g, in/g) with g = gcd."""
    g = calculate_gcd(in_rate, out_rate)
    return out_rate // g, in_rate // g


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------
class _Handle:
    _destroy = ""

    def __init__(self):
        self.h = _vp()

    def close(self):
        if getattr(self, "h", None):
            getattr(lib(), self._destroy)(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FreqShift(_Handle):
    """NCO frequency shift with one or more output shifts.

    ``FreqShift(fs, f)`` mirrors falcon_dsp_freq_shift::initialize(fs, f);
    ``FreqShift(fs, [f0, f1, ...])`` mirrors falcon_dsp_freq_shift_cuda's
    multi-channel initialize. ``apply`` returns one array per shift."""
    _destroy = "fdsp_freq_shift_destroy"

    def __init__(self, fs: int, shifts):
        super().__init__()
        self.single = np.isscalar(shifts)
        sh = [int(shifts)] if self.single else [int(s) for s in shifts]
        self.n_shifts = len(sh)
        arr = (C.c_int32 * len(sh))(*sh)
        check(lib().fdsp_freq_shift_create(fs, arr, len(sh), C.byref(self.h)))

    def reset_state(self):
        check(lib().fdsp_freq_shift_reset(self.h))

    def samples_handled(self, chan: int = 0) -> int:
        v = C.c_uint64()
        check(lib().fdsp_freq_shift_get_state(self.h, chan, C.byref(v)))
        return v.value

    def set_state(self, samples_handled: int):
        """Continue as if samples_handled samples had been processed (fdsp_freq_shift_set_state)."""
        check(lib().fdsp_freq_shift_set_state(self.h, int(samples_handled)))

    def apply_chunks_device(self, d_in: DeviceBuffer, chunk_len: int, n_chunks: int, d_out: DeviceBuffer,
                            out_stride: int, stream: Optional[Stream] = None, in_offset: int = 0):
        """n_chunks consecutive chunks in one launch (fdsp_freq_shift_apply_chunks_device)."""
        check(lib().fdsp_freq_shift_apply_chunks_device(self.h, d_in.at(in_offset), chunk_len, n_chunks, d_out.ptr,
                                                        out_stride, stream.h if stream else None))

    def apply(self, x):
        x = _cf32(x)
        outs = [np.empty(x.size, dtype=np.complex64) for _ in range(self.n_shifts)]
        ptrs = (_vp * self.n_shifts)(*[o.ctypes.data for o in outs])
        check(lib().fdsp_freq_shift_apply(self.h, _ptr(x), x.size, ptrs))
        return outs[0] if self.single else outs

    def apply_device(self, d_in: DeviceBuffer, n: int, d_out: DeviceBuffer, out_stride: int,
                     stream: Optional[Stream] = None, in_offset: int = 0):
        check(lib().fdsp_freq_shift_apply_device(self.h, d_in.at(in_offset), n, d_out.ptr, out_stride,
                                                 stream.h if stream else None))

    def apply_ci16(self, x):
        """falcon_dsp_freq_shift::apply(vector<complex<int16_t>>) (reference
        freq_shift.cc:133-162): (n, 2) int16 in, (n, 2) int16 out per shift."""
        x = _ci16(x)
        n = x.shape[0]
        outs = [np.empty((n, 2), dtype=np.int16) for _ in range(self.n_shifts)]
        ptrs = (_vp * self.n_shifts)(*[o.ctypes.data for o in outs])
        check(lib().fdsp_freq_shift_apply_ci16(self.h, _ptr(x), n, ptrs))
        return outs[0] if self.single else outs

    def apply_ci16_device(self, d_in: DeviceBuffer, n: int, d_out: DeviceBuffer, out_stride: int,
                          stream: Optional[Stream] = None, in_offset: int = 0):
        """Device buffers of CI16 samples (DeviceBuffer(count, CI16))."""
        check(lib().fdsp_freq_shift_apply_ci16_device(self.h, d_in.at(in_offset), n, d_out.ptr, out_stride,
                                                      stream.h if stream else None))


class PolyphaseResampler(_Handle):
    """falcon_dsp_polyphase_resampler: streaming upfirdn with real taps."""
    _destroy = "fdsp_resampler_destroy"

    def __init__(self, up: int, down: int, taps):
        super().__init__()
        t = _cf32(taps)
        self._taps = t
        self.up, self.down = int(up), int(down)
        check(lib().fdsp_resampler_create(up, down, _ptr(t), t.size, C.byref(self.h)))

    def reset_state(self):
        check(lib().fdsp_resampler_reset(self.h))

    def needed_out_count(self, in_count: int) -> int:
        v = C.c_uint32()
        check(lib().fdsp_resampler_needed_out_count(self.h, in_count, C.byref(v)))
        return v.value

    def coeffs_per_phase(self) -> int:
        v = C.c_uint32()
        check(lib().fdsp_resampler_coeffs_per_phase(self.h, C.byref(v)))
        return v.value

    def state(self):
        """(coeff_phase, xOffset, inputs consumed, outputs produced)."""
        ph, xo, it, ot = C.c_uint32(), C.c_uint32(), C.c_uint64(), C.c_uint64()
        check(lib().fdsp_resampler_get_state(self.h, C.byref(ph), C.byref(xo), C.byref(it), C.byref(ot)))
        return ph.value, xo.value, it.value, ot.value

    def apply(self, x) -> np.ndarray:
        x = _cf32(x)
        cap = self.needed_out_count(x.size)
        out = np.empty(max(cap, 1), dtype=np.complex64)
        n = C.c_size_t()
        check(lib().fdsp_resampler_apply(self.h, _ptr(x), x.size, _ptr(out), cap, C.byref(n)))
        return out[: n.value]

    def apply_device(self, d_in: DeviceBuffer, n: int, d_out: DeviceBuffer, stream: Optional[Stream] = None,
                     in_offset: int = 0) -> int:
        got = C.c_size_t()
        check(lib().fdsp_resampler_apply_device(self.h, d_in.at(in_offset), n, d_out.ptr, d_out.count,
                                                C.byref(got), stream.h if stream else None))
        return got.value

    def apply_chunks_device(self, d_in: DeviceBuffer, chunk_len: int, n_chunks: int, d_out: DeviceBuffer,
                            stream: Optional[Stream] = None, in_offset: int = 0) -> List[int]:
        """n_chunks consecutive chunks in one launch (fdsp_resampler_apply_chunks_device);
        the outputs land back to back, the per-chunk counts are returned."""
        counts = (C.c_size_t * max(n_chunks, 1))()
        check(lib().fdsp_resampler_apply_chunks_device(self.h, d_in.at(in_offset), chunk_len, n_chunks, d_out.ptr,
                                                       d_out.count, counts, stream.h if stream else None))
        return list(counts[:n_chunks])

    def set_state(self, coeff_phase: int, x_offset: int, history):
        """Restore the reference's coeff_phase, xOffset and state (the last
        coeffs_per_phase - 1 inputs, oldest first) (fdsp_resampler_set_state)."""
        h = _cf32(history)
        check(lib().fdsp_resampler_set_state(self.h, int(coeff_phase), int(x_offset), _ptr(h) if h.size else None,
                                             h.size))


class FirFilter(_Handle):
    """falcon_dsp_fir_filter: y[n] = sum_k h[k] x[n-k] with carried history."""
    _destroy = "fdsp_fir_destroy"

    def __init__(self, coeffs):
        super().__init__()
        c = _cf32(coeffs)
        self._c = c
        check(lib().fdsp_fir_create(_ptr(c), c.size, C.byref(self.h)))

    def reset_state(self):
        check(lib().fdsp_fir_reset(self.h))

    def apply(self, x) -> np.ndarray:
        x = _cf32(x)
        out = np.empty(x.size, dtype=np.complex64)
        check(lib().fdsp_fir_apply(self.h, _ptr(x), x.size, _ptr(out)))
        return out

    def apply_device(self, d_in: DeviceBuffer, n: int, d_out: DeviceBuffer, stream: Optional[Stream] = None):
        check(lib().fdsp_fir_apply_device(self.h, d_in.ptr, n, d_out.ptr, stream.h if stream else None))

    def apply_ci16(self, x) -> np.ndarray:
        """falcon_dsp_fir_filter::apply(vector<complex<int16_t>>) (reference
        fir_filter.cc:151-175): (n, 2) int16 in and out; the history is shared
        with apply()."""
        x = _ci16(x)
        out = np.empty_like(x)
        check(lib().fdsp_fir_apply_ci16(self.h, _ptr(x), x.shape[0], _ptr(out)))
        return out

    def apply_ci16_device(self, d_in: DeviceBuffer, n: int, d_out: DeviceBuffer, stream: Optional[Stream] = None):
        check(lib().fdsp_fir_apply_ci16_device(self.h, d_in.ptr, n, d_out.ptr, stream.h if stream else None))


class FirBank(_Handle):
    """Many independent FIR channels with real taps, one launch for all."""
    _destroy = "fdsp_fir_bank_destroy"

    def __init__(self, taps: np.ndarray):
        super().__init__()
        t = np.ascontiguousarray(np.asarray(taps, dtype=np.float32))
        if t.ndim != 2:
            raise ValueError("taps must be (n_chan, n_taps)")
        self.n_chan, self.n_taps = t.shape
        self._t = t
        check(lib().fdsp_fir_bank_create(self.n_chan, self.n_taps, _ptr(t), C.byref(self.h)))

    def reset_state(self):
        check(lib().fdsp_fir_bank_reset(self.h))

    def apply_device(self, d_in: DeviceBuffer, n: int, in_stride: int, d_out: DeviceBuffer, out_stride: int,
                     stream: Optional[Stream] = None):
        check(lib().fdsp_fir_bank_apply_device(self.h, d_in.ptr, n, in_stride, d_out.ptr, out_stride,
                                               stream.h if stream else None))

    def apply(self, x: np.ndarray) -> np.ndarray:
        """x: (n_chan, n) complex64 -> (n_chan, n)."""
        x = _cf32(x).reshape(self.n_chan, -1)
        n = x.shape[1]
        d_in = DeviceBuffer.from_array(x.ravel())
        d_out = DeviceBuffer(x.size)
        self.apply_device(d_in, n, n, d_out, n)
        device_sync()
        return d_out.download().reshape(self.n_chan, n)


class PolarDiscriminator(_Handle):
    """falcon_dsp_polar_discriminator: atan2 of x[n]*conj(x[n-1])."""
    _destroy = "fdsp_discriminator_destroy"

    def __init__(self):
        super().__init__()
        check(lib().fdsp_discriminator_create(C.byref(self.h)))

    def reset_state(self):
        check(lib().fdsp_discriminator_reset(self.h))

    def apply(self, x) -> np.ndarray:
        x = _cf32(x)
        out = np.empty(max(x.size, 1), dtype=np.float32)
        n = C.c_size_t()
        check(lib().fdsp_discriminator_apply(self.h, _ptr(x), x.size, _ptr(out), C.byref(n)))
        return out[: n.value]

    def apply_ci16(self, x) -> np.ndarray:
        """falcon_dsp_polar_discriminator::apply(vector<complex<int16_t>>)
        (reference polar_discriminator.cc:117-132); the carried sample is
        shared with apply()."""
        x = _ci16(x)
        out = np.empty(max(x.shape[0], 1), dtype=np.float32)
        n = C.c_size_t()
        check(lib().fdsp_discriminator_apply_ci16(self.h, _ptr(x), x.shape[0], _ptr(out), C.byref(n)))
        return out[: n.value]


class IirFilter(_Handle):
    """IIR filter y[n] = sum_k b[k] x[n-k] - sum_{j>=1} a[j] y[n-j]
    (falcon_dsp_iir_filter; reference hdr/transform/falcon_dsp_iir_filter.h:
    97-130, src/transform/falcon_dsp_iir_filter.cc:104-396). When a[0]'s real
    part is not 1 all coefficients are divided by a[0], as the reference does.
    The four apply forms share the stream state. int16 outputs are the
    reference's bit for bit; float outputs are too whenever the filter's
    parallel segments verify (first-order filters), else within float
    rounding unless set_exact(True) asks for the sequential exact form.
    int16 outputs are truncated toward zero and saturated. Limits:
    1 <= len(b) <= 256, 1 <= len(a) <= 17, a[0] != 0."""
    _destroy = "fdsp_iir_destroy"

    def __init__(self, b, a):
        super().__init__()
        b = np.ascontiguousarray(np.atleast_1d(b), dtype=np.complex64)
        a = np.ascontiguousarray(np.atleast_1d(a), dtype=np.complex64)
        check(lib().fdsp_iir_create(_ptr(b), b.size, _ptr(a), a.size, C.byref(self.h)))

    def reset_state(self):
        check(lib().fdsp_iir_reset(self.h))

    PLANS = {0: "fast", 1: "segments", 2: "serial"}

    @property
    def warm(self) -> int:
        """samples each segment's recurrence runs before its outputs"""
        w = C.c_uint32()
        check(lib().fdsp_iir_info(self.h, C.byref(w), None, None))
        return w.value

    @property
    def plan(self) -> str:
        """how a call with complex float outputs runs: "segments" (parallel,
        verified exact), "serial" (one sequential wave, exact) or "fast"
        (parallel, float rounding)"""
        p = C.c_uint32()
        check(lib().fdsp_iir_info(self.h, None, None, C.byref(p)))
        return self.PLANS[p.value]

    def set_exact(self, exact: Optional[bool] = None):
        """None (default): exact where parallel segments verify, else float
        rounding for float outputs and exact for int16 ones; True: always the
        reference's outputs bit for bit; False: always the unverified parallel
        segments"""
        check(lib().fdsp_iir_set_mode(self.h, 2 if exact is None else 1 if exact else 0))

    def apply(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.complex64)
        out = np.empty_like(x)
        check(lib().fdsp_iir_apply(self.h, _ptr(x), x.size, _ptr(out)))
        return out

    def apply_f32(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty_like(x)
        check(lib().fdsp_iir_apply_f32(self.h, _ptr(x), x.size, _ptr(out)))
        return out

    def apply_ci16(self, x) -> np.ndarray:
        """x: (n, 2) int16 (re, im) pairs; returns the same layout"""
        x = _ci16(x)
        out = np.empty_like(x)
        check(lib().fdsp_iir_apply_ci16(self.h, _ptr(x), x.shape[0], _ptr(out)))
        return out

    def apply_i16(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.int16)
        out = np.empty_like(x)
        check(lib().fdsp_iir_apply_i16(self.h, _ptr(x), x.size, _ptr(out)))
        return out

    def apply_device(self, d_in: DeviceBuffer, n: int, d_out: DeviceBuffer, stream: Optional[Stream] = None,
                     kind: str = "cf32"):
        """kind: "cf32", "f32", "ci16" or "i16" (both buffers in that layout)"""
        fn = {"cf32": "fdsp_iir_apply_device", "f32": "fdsp_iir_apply_f32_device",
              "ci16": "fdsp_iir_apply_ci16_device", "i16": "fdsp_iir_apply_i16_device"}[kind]
        check(getattr(lib(), fn)(self.h, d_in.ptr, n, d_out.ptr, stream.h if stream else None))

    def apply_device_noalloc(self, d_in: DeviceBuffer, n: int, d_out: DeviceBuffer, stream: Optional[Stream] = None):
        """Complex float buffers; the kernels' work memory is the library's static
        per-device array (fdsp_iir_apply_device_noalloc): nothing is allocated,
        and overlapping calls on one device race."""
        check(lib().fdsp_iir_apply_device_noalloc(self.h, d_in.ptr, n, d_out.ptr, stream.h if stream else None))


class IirBank(_Handle):
    """n_streams independent falcon_dsp_iir_filter streams sharing (b, a),
    each bit-identical to the reference's sequential loop (one stream per GPU
    lane; fdsp_iir_bank_*). apply takes an (n_streams, n) array."""
    _destroy = "fdsp_iir_bank_destroy"

    def __init__(self, b, a, n_streams: int):
        super().__init__()
        self._b = np.ascontiguousarray(b, dtype=np.complex64)
        self._a = np.ascontiguousarray(a, dtype=np.complex64)
        self.n_streams = int(n_streams)
        check(lib().fdsp_iir_bank_create(_ptr(self._b), self._b.size, _ptr(self._a), self._a.size,
                                         self.n_streams, C.byref(self.h)))

    def reset_state(self):
        check(lib().fdsp_iir_bank_reset(self.h))

    def apply(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.complex64)
        if x.ndim != 2 or x.shape[0] != self.n_streams:
            raise ValueError(f"need an ({self.n_streams}, n) array")
        out = np.empty_like(x)
        check(lib().fdsp_iir_bank_apply(self.h, _ptr(x), x.shape[1], _ptr(out)))
        return out

    def apply_ci16(self, x) -> np.ndarray:
        """x: (n_streams, n, 2) int16 (re, im) pairs; returns the same layout"""
        x = np.ascontiguousarray(x, dtype=np.int16)
        if x.ndim != 3 or x.shape[0] != self.n_streams or x.shape[2] != 2:
            raise ValueError(f"need an ({self.n_streams}, n, 2) int16 array")
        out = np.empty_like(x)
        check(lib().fdsp_iir_bank_apply_ci16(self.h, _ptr(x), x.shape[1], _ptr(out)))
        return out

    def apply_device(self, d_in: DeviceBuffer, n: int, in_stride: int, d_out: DeviceBuffer, out_stride: int,
                     stream: Optional[Stream] = None, ci16: bool = False):
        fn = lib().fdsp_iir_bank_apply_ci16_device if ci16 else lib().fdsp_iir_bank_apply_device
        check(fn(self.h, d_in.ptr, n, in_stride, d_out.ptr, out_stride, stream.h if stream else None))


def iir_filter(b, a, x) -> np.ndarray:
    """One-shot IIR filter (reference iir_filter.cc:85-100): a fresh filter over x."""
    x = np.asarray(x)
    f = IirFilter(b, a)
    if x.dtype == np.int16 and x.ndim == 2:
        return f.apply_ci16(x)
    return f.apply(x)


class Iir1(_Handle):
    """First-order IIR on real samples, y[n] = b0 x[n] - a1 y[n-1]
    (falcon_dsp_iir_filter with b = [b0], a = [1, a1]; reference
    iir_filter.cc:104-396); outputs bit-identical to the reference's loop."""
    _destroy = "fdsp_iir1_destroy"

    def __init__(self, b0: float, a1: float):
        super().__init__()
        check(lib().fdsp_iir1_create(float(b0), float(a1), C.byref(self.h)))

    def reset_state(self):
        check(lib().fdsp_iir1_reset(self.h))

    def apply(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty_like(x)
        check(lib().fdsp_iir1_apply(self.h, _ptr(x), x.size, _ptr(out)))
        return out

    def apply_device(self, d_in: DeviceBuffer, n: int, d_out: DeviceBuffer, stream: Optional[Stream] = None):
        check(lib().fdsp_iir1_apply_device(self.h, d_in.ptr, n, d_out.ptr, stream.h if stream else None))


def fm_deemphasis(tau: float):
    """(b0, a1) of the FM de-emphasis for time constant tau (reference fm_demod.cc:146-165)."""
    b0, a1 = C.c_float(), C.c_float()
    check(lib().fdsp_fm_deemphasis(float(tau), C.byref(b0), C.byref(a1)))
    return b0.value, a1.value


FILTER_REMEZ, FILTER_FIRLS, FILTER_TAPS_OPTIMAL = 0, 1, -1


def resample_fir_params(in_rate: int, out_rate: int, taps: int = 64, source: int = FILTER_REMEZ):
    """The predefined anti-alias filter for in_rate -> out_rate sps
    (get_resample_fir_params; reference src/transform/falcon_dsp_fir_filter.cc:
    111-130): (up, down, complex64 taps). taps: 32, 64, 128, 256 or
    FILTER_TAPS_OPTIMAL; source: FILTER_REMEZ or FILTER_FIRLS. Raises
    FdspError when the table has no such filter."""
    up, down, n = C.c_uint32(), C.c_uint32(), C.c_size_t()
    buf = np.empty(2048, np.complex64)
    check(lib().fdsp_resample_fir_params(int(in_rate), int(out_rate), int(taps), int(source), C.byref(up),
                                         C.byref(down), _ptr(buf), buf.size, C.byref(n)))
    return up.value, down.value, buf[:n.value].copy()


class FmDemodulator(_Handle):
    """falcon_dsp_fm_demodulator::demod_mono (reference src/signals/falcon_dsp_fm_demod.cc:98-271):
    shift, resample to the FM band, discriminate, de-emphasise, resample to
    audio, scale to a peak of 10000 and convert to int16. iso and mono are
    the two resamplers' (up, down, taps); by default the predefined 64-tap
    REMEZ filters for fs -> 228,000 and 228,000 -> 45,600 sps, as the
    reference takes them."""
    _destroy = "fdsp_fm_demod_destroy"
    AMERICAS_TAU = 75e-6
    EUROPE_ASIA_TAU = 50e-6
    FM_RATE, AUDIO_RATE = 228_000, 45_600

    def __init__(self, fs: int, offset_hz: int, iso=None, mono=None, tau: float = AMERICAS_TAU):
        super().__init__()
        if iso is None:
            iso = resample_fir_params(fs, self.FM_RATE)
        if mono is None:
            mono = resample_fir_params(self.FM_RATE, self.AUDIO_RATE)
        self._iso_t, self._mono_t = _cf32(iso[2]), _cf32(mono[2])
        check(lib().fdsp_fm_demod_create(fs, int(offset_hz), float(tau), iso[0], iso[1], _ptr(self._iso_t),
                                         self._iso_t.size, mono[0], mono[1], _ptr(self._mono_t), self._mono_t.size,
                                         C.byref(self.h)))

    def reset_state(self):
        check(lib().fdsp_fm_demod_reset(self.h))

    def out_count(self, n: int) -> int:
        v = C.c_size_t()
        check(lib().fdsp_fm_demod_out_count(self.h, n, C.byref(v)))
        return v.value

    def demod_mono(self, x, audio: bool = False):
        """-> int16 samples, or (int16, float samples before scaling) with audio=True."""
        x = _cf32(x)
        cap = max(self.out_count(x.size), 1)
        out = np.empty(cap, dtype=np.int16)
        a = np.empty(cap, dtype=np.float32) if audio else None
        n = C.c_size_t()
        check(lib().fdsp_fm_demod_apply(self.h, _ptr(x), x.size, _ptr(out), _ptr(a) if audio else None, cap,
                                        C.byref(n)))
        return (out[: n.value], a[: n.value]) if audio else out[: n.value]


class MultiRateChannelizer(_Handle):
    """falcon_dsp_multi_rate_channelizer{,_cuda}: per channel, NCO shift then
    polyphase resample.

    ``channels``: sequence of dicts with the fields of
    multi_rate_channelizer_channel_s (output_sample_rate_in_sps,
    freq_shift_in_hz, up_rate, down_rate, taps) and optionally
    ``tail="discriminator"``. ``batched=True`` selects the fused one-launch
    kernel; ``batched=False`` the reference-structured per-channel form.
    ``arith``: "f32" (the default: FP32 arithmetic throughout) or "bf16x2"
    (the batched form's opt-in reduced-precision matrix products)."""
    _destroy = "fdsp_channelizer_destroy"

    def __init__(self, fs: int, channels: Sequence[dict], batched: bool = True, arith: str = "f32"):
        super().__init__()
        self.fs = int(fs)
        self.n = len(channels)
        self._taps = [_cf32(c["taps"]) for c in channels]
        self.tails = [1 if c.get("tail") in ("discriminator", N.FDSP_TAIL_DISCRIMINATOR) else 0 for c in channels]
        descs = (N.ChannelDesc * self.n)()
        for i, c in enumerate(channels):
            d = descs[i]
            d.output_sample_rate_in_sps = int(c.get("output_sample_rate_in_sps", 0))
            d.freq_shift_in_hz = int(c["freq_shift_in_hz"])
            d.up_rate = int(c["up_rate"])
            d.down_rate = int(c["down_rate"])
            d.taps = self._taps[i].ctypes.data
            d.n_taps = self._taps[i].size
            d.tail = self.tails[i]
        if arith not in ("f32", "bf16x2"):
            raise ValueError(f"arith must be 'f32' or 'bf16x2', not {arith!r}")
        if not batched:
            self.mode = N.FDSP_CHAN_PLAIN
        else:
            self.mode = N.FDSP_CHAN_BATCHED if arith == "f32" else N.FDSP_CHAN_BATCHED_BF16X2
        check(lib().fdsp_channelizer_create(self.fs, descs, self.n, self.mode, C.byref(self.h)))
        self.ups = [int(c["up_rate"]) for c in channels]
        self.downs = [int(c["down_rate"]) for c in channels]

    def reset_state(self):
        check(lib().fdsp_channelizer_reset(self.h))

    def needed_out_count(self, chan: int, in_count: int) -> int:
        v = C.c_uint32()
        check(lib().fdsp_channelizer_needed_out_count(self.h, chan, in_count, C.byref(v)))
        return v.value

    def state(self, chan: int):
        """(nco samples handled, coeff_phase, xOffset, inputs consumed, outputs produced)."""
        k, ph, xo, it, ot = C.c_uint64(), C.c_uint32(), C.c_uint32(), C.c_uint64(), C.c_uint64()
        check(lib().fdsp_channelizer_get_state(self.h, chan, C.byref(k), C.byref(ph), C.byref(xo), C.byref(it),
                                               C.byref(ot)))
        return k.value, ph.value, xo.value, it.value, ot.value

    def set_tiling(self, segment_len: int, block_threads: int):
        check(lib().fdsp_channelizer_set_tiling(self.h, segment_len, block_threads))

    def get_tiling(self):
        s, b = C.c_uint32(), C.c_uint32()
        check(lib().fdsp_channelizer_get_tiling(self.h, C.byref(s), C.byref(b)))
        return s.value, b.value

    def set_kernel(self, kernel: str):
        """"segment" (FDSP_KERNEL_SEGMENT_F32) or "groups" (FDSP_KERNEL_GROUPS_F32)."""
        k = {"segment": N.FDSP_KERNEL_SEGMENT_F32, "groups": N.FDSP_KERNEL_GROUPS_F32}[kernel]
        check(lib().fdsp_channelizer_set_kernel(self.h, k))

    def kernel_info(self) -> dict:
        """Which kernel the batched form launches (fdsp_channelizer_kernel_info)."""
        k, g, m = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().fdsp_channelizer_kernel_info(self.h, C.byref(k), C.byref(g), C.byref(m)))
        return {"kind": k.value, "kernel": N.KERNEL_NAMES[k.value], "groups": g.value, "mfma_groups": m.value}

    def autotune(self, d_in: DeviceBuffer, chunk_len: int, reps: int = 10, in_offset: int = 0) -> List[dict]:
        """Time every candidate work split on a device chunk and keep the
        fastest (fdsp_channelizer_autotune); returns the trials. A pick stored
        in the tuning cache (FDSP_TUNE_CACHE) comes back as one trial with
        "cached" set."""
        cap = 32
        tr = (N.TilingTrial * cap)()
        k = C.c_size_t()
        check(lib().fdsp_channelizer_autotune(self.h, d_in.at(in_offset), chunk_len, reps, tr, cap, C.byref(k)))
        return [{"segment_len": t.segment_len, "block_threads": t.block_threads, "us": round(t.us, 2),
                 "measured_weights": bool(t.flags & 1), "fused": bool(t.flags & 2),
                 "segment_kernel": bool(t.flags & 4), "cached": bool(t.flags & 8)} for t in tr[: k.value]]

    def set_fusion(self, on: bool):
        """Phase grouping: bf16-plane groups share staged spans (fdsp_channelizer_set_fusion)."""
        check(lib().fdsp_channelizer_set_fusion(self.h, 1 if on else 0))

    def get_fusion(self):
        """(on, number of groups in the fused record)."""
        on, k = C.c_int(), C.c_uint32()
        check(lib().fdsp_channelizer_get_fusion(self.h, C.byref(on), C.byref(k)))
        return bool(on.value), k.value

    def plan_info(self, chunk_len: int) -> dict:
        """Work split of the batched kernel for a chunk length (diagnostics)."""
        v = [C.c_uint32() for _ in range(4)]
        check(lib().fdsp_channelizer_plan_info(self.h, chunk_len, *[C.byref(x) for x in v]))
        return dict(zip(("blocks", "capacity", "groups", "lds_bytes"), (x.value for x in v)))

    def apply(self, x) -> List[np.ndarray]:
        x = _cf32(x)
        caps = [self.needed_out_count(i, x.size) for i in range(self.n)]
        outs = [np.empty(max(c, 1), dtype=np.float32 if self.tails[i] else np.complex64)
                for i, c in enumerate(caps)]
        op = (_vp * self.n)(*[o.ctypes.data for o in outs])
        cp = (C.c_size_t * self.n)(*caps)
        counts = (C.c_size_t * self.n)()
        check(lib().fdsp_channelizer_apply(self.h, _ptr(x), x.size, op, cp, counts))
        return [outs[i][: counts[i]] for i in range(self.n)]

    def max_out(self, chunk: int) -> List[int]:
        """Upper bound of outputs per channel for any chunk of ``chunk`` inputs."""
        return [-(-chunk * u // d) for u, d in zip(self.ups, self.downs)]

    def apply_device(self, d_in: DeviceBuffer, n: int, d_outs: Sequence[DeviceBuffer],
                     stream: Optional[Stream] = None, in_offset: int = 0) -> List[int]:
        op = (_vp * self.n)(*[b.ptr.value for b in d_outs])
        cp = (C.c_size_t * self.n)(*[b.count for b in d_outs])
        counts = (C.c_size_t * self.n)()
        check(lib().fdsp_channelizer_apply_device(self.h, d_in.at(in_offset), n, op, cp, counts,
                                                  stream.h if stream else None))
        return list(counts)

    # graph mode
    def graph_create(self, chunk_len: int, d_in: DeviceBuffer, d_outs: Sequence[DeviceBuffer]):
        op = (_vp * self.n)(*[b.ptr.value for b in d_outs])
        cp = (C.c_size_t * self.n)(*[b.count for b in d_outs])
        check(lib().fdsp_channelizer_graph_create(self.h, chunk_len, d_in.ptr, op, cp))

    def graph_launch(self, stream: Optional[Stream] = None):
        check(lib().fdsp_channelizer_graph_launch(self.h, stream.h if stream else None))

    def graph_counts(self) -> List[int]:
        counts = (C.c_size_t * self.n)()
        check(lib().fdsp_channelizer_graph_counts(self.h, counts))
        return list(counts)

    def graph_destroy(self):
        check(lib().fdsp_channelizer_graph_destroy(self.h))


# ---------------------------------------------------------------------------
# serving
# ---------------------------------------------------------------------------
class Server(_Handle):
    """Continuous-batching server of channelizer streams (fdsp_serve_*).

    ``admit(channel)`` takes one channel dict (as MultiRateChannelizer's,
    optionally with ``tail="discriminator"``) and returns a stream id, or
    None when the stream limit or the device-memory budget is reached.
    ``submit(stream, x)`` queues a chunk and returns a ticket; ``poll`` and
    ``result`` collect it. The server keeps each request's input and output
    arrays alive until the ticket is collected."""
    _destroy = "fdsp_serve_destroy"

    def __init__(self, fs: int, max_streams: int = 200, max_chunk: int = 1 << 18, max_batch: int = 1 << 22,
                 hbm_budget: int = 64 << 30, state_bytes: int = 0):
        super().__init__()
        cfg = N.ServeConfig(int(fs), int(max_streams), int(max_chunk), int(max_batch), int(hbm_budget),
                            int(state_bytes))
        check(lib().fdsp_serve_create(C.byref(cfg), C.byref(self.h)))
        self.fs = int(fs)
        self._tail = {}
        self._live = {}

    def admit(self, channel: dict) -> Optional[int]:
        taps = _cf32(channel["taps"])
        tail = 1 if channel.get("tail") in ("discriminator", N.FDSP_TAIL_DISCRIMINATOR) else 0
        d = N.ChannelDesc(int(channel.get("output_sample_rate_in_sps", 0)), int(channel["freq_shift_in_hz"]),
                          int(channel["up_rate"]), int(channel["down_rate"]), taps.ctypes.data, taps.size, tail)
        sid = C.c_uint32()
        rc = lib().fdsp_serve_admit(self.h, C.byref(d), C.byref(sid))
        if rc == N.FDSP_EAGAIN:
            return None
        check(rc)
        self._tail[sid.value] = tail
        return sid.value

    def max_out(self, stream: int, n: int) -> int:
        v = C.c_size_t()
        check(lib().fdsp_serve_max_out(self.h, stream, n, C.byref(v)))
        return v.value

    def submit(self, stream: int, x) -> int:
        x = _cf32(x)
        out = np.empty(self.max_out(stream, x.size), dtype=np.float32 if self._tail[stream] else np.complex64)
        t = C.c_uint64()
        check(lib().fdsp_serve_submit(self.h, stream, _ptr(x), x.size, _ptr(out), out.size, C.byref(t)))
        self._live[t.value] = (x, out)
        return t.value

    def poll(self, ticket: int, wait_ms: int = 0):
        """(output array, fdsp_serve_result) when done, else None."""
        res = N.ServeResult()
        rc = lib().fdsp_serve_poll(self.h, ticket, int(wait_ms), C.byref(res))
        if rc == N.FDSP_EAGAIN:
            return None
        entry = self._live.pop(ticket, None)  # a finished ticket is consumed, failed or not
        check(rc)
        return entry[1][: res.n_out], res

    def result(self, ticket: int):
        """Wait for a ticket: (output array, fdsp_serve_result)."""
        return self.poll(ticket, -1)

    def release(self, stream: int):
        check(lib().fdsp_serve_release(self.h, stream))
        self._tail.pop(stream, None)

    def info(self) -> dict:
        st = N.ServeStats()
        check(lib().fdsp_serve_info(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in N.ServeStats._fields_}


# ---------------------------------------------------------------------------
# free functions (one-shot wrappers, as in the reference)
# ---------------------------------------------------------------------------
def freq_shift(in_sample_rate_in_sps: int, x, freq_shift_in_hz):
    """freq_shift / freq_shift_cuda (reference freq_shift.h:85, freq_shift_cuda.h:140-156)."""
    return FreqShift(in_sample_rate_in_sps, freq_shift_in_hz).apply(x)


def resample(in_sample_rate_in_sps: int, x, filter_coeffs, out_sample_rate_in_sps: int):
    """resample / resample_cuda (reference src/transform/falcon_dsp_resample.cc:126-137):
    up/down from rat_approx(out/in, 1024), then one apply."""
    up, down = rat_approx(out_sample_rate_in_sps / in_sample_rate_in_sps, 1024)
    if up <= 0 or down <= 0:
        return None
    return PolyphaseResampler(up, down, filter_coeffs).apply(x)


upfirdn = resample


def fir_filter(coeffs, x):
    """fir_filter / fir_filter_cuda (reference fir_filter.cc:92-97)."""
    return FirFilter(coeffs).apply(x)


def polar_discriminator(x):
    """polar_discriminator (reference polar_discriminator.cc:91-96)."""
    return PolarDiscriminator().apply(x)
