"""ctypes access to the CPU checkers (test infrastructure only).

* ``liboracle.so``  - the C restatement of the reference algorithms
  (``oracle/fdsp_oracle.c``; each function cites the reference file:line it
  follows).
* ``_ref/libfdsp_ref.so`` - the reference's own CPU classes compiled from the
  reference sources (``oracle/Makefile``); present only where it was built.
* FP64 closed forms (numpy) used as a third, independent check.

Only tests, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module. The product package never does.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libfdsp_ref.so")

_f32p = C.POINTER(C.c_float)
_vp = C.c_void_p


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(_f32p)


_i16p = C.POINTER(C.c_int16)


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(_i16p)


def as_ci16(x) -> np.ndarray:
    """(n, 2) int16 array of (re, im) pairs, the layout of std::complex<int16_t>."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.int16))
    assert a.ndim == 2 and a.shape[1] == 2
    return a


def as_cf32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.complex64))


# --------------------------------------------------------------------------
# the C restatement
# --------------------------------------------------------------------------
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle oracle`")
        L = C.CDLL(ORACLE_SO)
        L.orc_gcd.restype = C.c_uint64
        L.orc_gcd.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_lcm.restype = C.c_uint64
        L.orc_lcm.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_rat_approx.argtypes = [C.c_double, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_freq_shift_params.restype = C.c_int
        L.orc_freq_shift_params.argtypes = [C.c_uint32, C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
        for name in ("freq_shift", "resampler", "fir", "discrim", "channelizer", "iir", "fm"):
            getattr(L, f"orc_sizeof_{name}").restype = C.c_size_t
        L.orc_freq_shift_init.argtypes = [_vp, C.c_uint32, C.c_int32]
        L.orc_freq_shift_reset.argtypes = [_vp]
        L.orc_freq_shift_apply.restype = C.c_size_t
        L.orc_freq_shift_apply.argtypes = [_vp, _f32p, C.c_size_t, _f32p]
        L.orc_resampler_init.argtypes = [_vp, C.c_uint32, C.c_uint32, _f32p, C.c_size_t]
        L.orc_resampler_free.argtypes = [_vp]
        L.orc_resampler_reset.argtypes = [_vp]
        L.orc_resampler_needed_out_count.restype = C.c_uint32
        L.orc_resampler_needed_out_count.argtypes = [_vp, C.c_uint32]
        L.orc_resampler_apply.restype = C.c_size_t
        L.orc_resampler_apply.argtypes = [_vp, _f32p, C.c_size_t, _f32p]
        L.orc_fir_init.argtypes = [_vp, _f32p, C.c_size_t]
        L.orc_fir_free.argtypes = [_vp]
        L.orc_fir_reset.argtypes = [_vp]
        L.orc_fir_apply.restype = C.c_size_t
        L.orc_fir_apply.argtypes = [_vp, _f32p, C.c_size_t, _f32p]
        L.orc_discrim_reset.argtypes = [_vp]
        L.orc_discrim_apply.restype = C.c_size_t
        L.orc_discrim_apply.argtypes = [_vp, _f32p, C.c_size_t, _f32p]
        L.orc_iir_init.argtypes = [_vp, _f32p, C.c_size_t, _f32p, C.c_size_t]
        L.orc_iir_free.argtypes = [_vp]
        L.orc_iir_reset.argtypes = [_vp]
        for name in ("orc_iir_apply", "orc_iir_apply_f32"):
            getattr(L, name).restype = C.c_size_t
            getattr(L, name).argtypes = [_vp, _f32p, C.c_size_t, _f32p]
        for name in ("orc_iir_apply_ci16", "orc_iir_apply_i16"):
            getattr(L, name).restype = C.c_size_t
            getattr(L, name).argtypes = [_vp, _i16p, C.c_size_t, _i16p]
        L.orc_fm_deemphasis.argtypes = [C.c_double, _f32p, _f32p]
        L.orc_fm_init.argtypes = [_vp, C.c_uint32, C.c_int32, C.c_double, C.c_uint32, C.c_uint32, _f32p, C.c_size_t,
                                  C.c_uint32, C.c_uint32, _f32p, C.c_size_t]
        L.orc_fm_free.argtypes = [_vp]
        L.orc_fm_reset.argtypes = [_vp]
        L.orc_fm_out_cap.restype = C.c_size_t
        L.orc_fm_out_cap.argtypes = [_vp, C.c_size_t]
        L.orc_fm_demod.restype = C.c_size_t
        L.orc_fm_demod.argtypes = [_vp, _f32p, C.c_size_t, _i16p, _f32p]
        L.orc_i16.restype = C.c_int16
        L.orc_i16.argtypes = [C.c_float]
        for name, outp in (("freq_shift", _i16p), ("fir", _i16p), ("discrim", _f32p)):
            f = getattr(L, f"orc_{name}_apply_ci16")
            f.restype = C.c_size_t
            f.argtypes = [_vp, _i16p, C.c_size_t, outp]
        L.orc_channelizer_init.argtypes = [_vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_int64),
                                           C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                           C.POINTER(_f32p), C.POINTER(C.c_uint32)]
        L.orc_channelizer_free.argtypes = [_vp]
        L.orc_channelizer_reset.argtypes = [_vp]
        L.orc_channelizer_needed_out_count.restype = C.c_uint32
        L.orc_channelizer_needed_out_count.argtypes = [_vp, C.c_uint32, C.c_uint32]
        L.orc_channelizer_apply.argtypes = [_vp, _f32p, C.c_size_t, C.POINTER(_f32p), C.POINTER(C.c_size_t)]
        L.orc_channelizer_state.argtypes = [_vp, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_uint32),
                                            C.POINTER(C.c_uint32)]
        _lib = L
    return _lib


def gcd(a, b):
    return int(lib().orc_gcd(a, b))


def lcm(a, b):
    return int(lib().orc_lcm(a, b))


def rat_approx(f, md):
    n, d = C.c_int64(), C.c_int64()
    lib().orc_rat_approx(f, md, C.byref(n), C.byref(d))
    return n.value, d.value


def freq_shift_params(fs, f):
    r, w = C.c_uint32(), C.c_double()
    if lib().orc_freq_shift_params(fs, f, C.byref(r), C.byref(w)) != 0:
        raise ValueError("invalid frequency-shift parameters")
    return r.value, w.value


class _Obj:
    _kind = ""

    def __init__(self):
        self._buf = C.create_string_buffer(getattr(lib(), f"orc_sizeof_{self._kind}")())
        self.p = C.cast(self._buf, _vp)


class FreqShift(_Obj):
    """falcon_dsp_freq_shift (src/transform/falcon_dsp_freq_shift.cc:107-192)."""
    _kind = "freq_shift"

    def __init__(self, fs, f):
        super().__init__()
        if lib().orc_freq_shift_init(self.p, fs, f) != 0:
            raise ValueError("init failed")

    def reset(self):
        lib().orc_freq_shift_reset(self.p)

    @property
    def samples_handled(self):
        # first field after the int flag (padded to 8): the double counter
        return C.cast(self._buf, C.POINTER(C.c_double))[1]

    def apply(self, x):
        x = as_cf32(x)
        out = np.empty_like(x)
        n = lib().orc_freq_shift_apply(self.p, _fptr(x.view(np.float32)), x.size, _fptr(out.view(np.float32)))
        return out[:n]

    def apply_ci16(self, x):
        """freq_shift.cc:133-162 (complex int16 in and out)."""
        x = as_ci16(x)
        out = np.empty_like(x)
        n = lib().orc_freq_shift_apply_ci16(self.p, _iptr(x), x.shape[0], _iptr(out))
        return out[:n]


class Resampler(_Obj):
    """falcon_dsp_polyphase_resampler (src/transform/falcon_dsp_polyphase_resampler.cc:114-346)."""
    _kind = "resampler"

    def __init__(self, up, down, taps):
        super().__init__()
        t = as_cf32(taps)
        self._taps = t
        if lib().orc_resampler_init(self.p, up, down, _fptr(t.view(np.float32)), t.size) != 0:
            raise ValueError("init failed")
        self.up, self.down = up, down

    def __del__(self):
        try:
            lib().orc_resampler_free(self.p)
        except Exception:
            pass

    def reset(self):
        lib().orc_resampler_reset(self.p)

    def needed_out_count(self, n):
        return int(lib().orc_resampler_needed_out_count(self.p, n))

    @property
    def state_fields(self):
        """(coeff_phase, x_offset) as carried between chunks."""
        u32 = np.frombuffer(self._buf, dtype=np.uint32)
        # layout: int initialized; u32 up, down, padded, cpp; [pad]; ptr bank; ptr state; u32 phase; u32 xoff
        return int(u32[10]), int(u32[11])

    def apply(self, x):
        x = as_cf32(x)
        out = np.empty(self.needed_out_count(x.size) + 1, dtype=np.complex64)
        n = lib().orc_resampler_apply(self.p, _fptr(x.view(np.float32)), x.size, _fptr(out.view(np.float32)))
        return out[:n]


class Fir(_Obj):
    """falcon_dsp_fir_filter (src/transform/falcon_dsp_fir_filter.cc:136-284)."""
    _kind = "fir"

    def __init__(self, coeffs):
        super().__init__()
        c = as_cf32(coeffs)
        self._c = c
        if lib().orc_fir_init(self.p, _fptr(c.view(np.float32)), c.size) != 0:
            raise ValueError("init failed")

    def __del__(self):
        try:
            lib().orc_fir_free(self.p)
        except Exception:
            pass

    def reset(self):
        lib().orc_fir_reset(self.p)

    def apply(self, x):
        x = as_cf32(x)
        out = np.empty_like(x)
        n = lib().orc_fir_apply(self.p, _fptr(x.view(np.float32)), x.size, _fptr(out.view(np.float32)))
        return out[:n]

    def apply_ci16(self, x):
        """fir_filter.cc:151-175 (complex int16 in and out)."""
        x = as_ci16(x)
        out = np.empty_like(x)
        n = lib().orc_fir_apply_ci16(self.p, _iptr(x), x.shape[0], _iptr(out))
        return out[:n]


class Discriminator(_Obj):
    """falcon_dsp_polar_discriminator (src/transform/falcon_dsp_polar_discriminator.cc:103-171)."""
    _kind = "discrim"

    def __init__(self):
        super().__init__()
        lib().orc_discrim_reset(self.p)

    def reset(self):
        lib().orc_discrim_reset(self.p)

    def apply(self, x):
        x = as_cf32(x)
        out = np.empty(x.size, dtype=np.float32)
        n = lib().orc_discrim_apply(self.p, _fptr(x.view(np.float32)), x.size, _fptr(out))
        return out[:n]

    def apply_ci16(self, x):
        """polar_discriminator.cc:117-132 (complex int16 in)."""
        x = as_ci16(x)
        out = np.empty(x.shape[0], dtype=np.float32)
        n = lib().orc_discrim_apply_ci16(self.p, _iptr(x), x.shape[0], _fptr(out))
        return out[:n]


class Iir(_Obj):
    """falcon_dsp_iir_filter (src/transform/falcon_dsp_iir_filter.cc:104-396)."""
    _kind = "iir"

    def __init__(self, b, a):
        super().__init__()
        self._b, self._a = as_cf32(b), as_cf32(a)
        if lib().orc_iir_init(self.p, _fptr(self._b.view(np.float32)), self._b.size,
                              _fptr(self._a.view(np.float32)), self._a.size) != 0:
            raise ValueError("init failed")

    def __del__(self):
        try:
            lib().orc_iir_free(self.p)
        except Exception:
            pass

    def reset(self):
        lib().orc_iir_reset(self.p)

    def apply(self, x):
        x = as_cf32(x)
        out = np.empty_like(x)
        n = lib().orc_iir_apply(self.p, _fptr(x.view(np.float32)), x.size, _fptr(out.view(np.float32)))
        return out[:n]

    def apply_f32(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty_like(x)
        n = lib().orc_iir_apply_f32(self.p, _fptr(x), x.size, _fptr(out))
        return out[:n]

    def apply_ci16(self, x):
        x = as_ci16(x)
        out = np.empty_like(x)
        n = lib().orc_iir_apply_ci16(self.p, _iptr(x), x.shape[0], _iptr(out))
        return out[:n]

    def apply_i16(self, x):
        x = np.ascontiguousarray(x, dtype=np.int16)
        out = np.empty_like(x)
        n = lib().orc_iir_apply_i16(self.p, _iptr(x), x.size, _iptr(out))
        return out[:n]


def fm_deemphasis(tau):
    """(b0, a1) of the single-pole de-emphasis (src/signals/falcon_dsp_fm_demod.cc:146-165)."""
    b0, a1 = C.c_float(), C.c_float()
    lib().orc_fm_deemphasis(float(tau), C.byref(b0), C.byref(a1))
    return b0.value, a1.value


class Fm(_Obj):
    """falcon_dsp_fm_demodulator::demod_mono (src/signals/falcon_dsp_fm_demod.cc:98-271)
    with the two resamplers' (up, down, taps) given."""
    _kind = "fm"

    def __init__(self, fs, offset_hz, tau, iso, mono):
        super().__init__()
        self._iso_t, self._mono_t = as_cf32(iso[2]), as_cf32(mono[2])
        if lib().orc_fm_init(self.p, fs, offset_hz, float(tau), iso[0], iso[1], _fptr(self._iso_t.view(np.float32)),
                             self._iso_t.size, mono[0], mono[1], _fptr(self._mono_t.view(np.float32)),
                             self._mono_t.size) != 0:
            raise ValueError("init failed")

    def __del__(self):
        try:
            lib().orc_fm_free(self.p)
        except Exception:
            pass

    def reset(self):
        lib().orc_fm_reset(self.p)

    def demod(self, x):
        """-> (int16 audio, float audio before scaling)."""
        x = as_cf32(x)
        cap = int(lib().orc_fm_out_cap(self.p, x.size))
        out = np.empty(cap, dtype=np.int16)
        audio = np.empty(cap, dtype=np.float32)
        n = lib().orc_fm_demod(self.p, _fptr(x.view(np.float32)), x.size, _iptr(out), _fptr(audio))
        return out[:n], audio[:n]


class Channelizer(_Obj):
    """falcon_dsp_multi_rate_channelizer (src/transform/falcon_dsp_multi_rate_channelizer.cc:81-173).

    ``channels`` is a list of dicts with keys freq_shift_in_hz, up_rate,
    down_rate, taps (complex64 array)."""
    _kind = "channelizer"

    def __init__(self, fs, channels):
        super().__init__()
        n = len(channels)
        self.n = n
        self._taps = [as_cf32(c["taps"]) for c in channels]
        shifts = (C.c_int64 * n)(*[int(c["freq_shift_in_hz"]) for c in channels])
        ups = (C.c_uint32 * n)(*[int(c["up_rate"]) for c in channels])
        downs = (C.c_uint32 * n)(*[int(c["down_rate"]) for c in channels])
        tp = (_f32p * n)(*[_fptr(t.view(np.float32)) for t in self._taps])
        nt = (C.c_uint32 * n)(*[t.size for t in self._taps])
        if lib().orc_channelizer_init(self.p, fs, n, shifts, ups, downs, tp, nt) != 0:
            raise ValueError("init failed")

    def __del__(self):
        try:
            lib().orc_channelizer_free(self.p)
        except Exception:
            pass

    def reset(self):
        lib().orc_channelizer_reset(self.p)

    def state(self, chan):
        """(NCO samples handled, coeff_phase, xOffset) of channel ``chan``."""
        s, ph, xo = C.c_double(), C.c_uint32(), C.c_uint32()
        if lib().orc_channelizer_state(self.p, chan, C.byref(s), C.byref(ph), C.byref(xo)) != 0:
            raise IndexError(chan)
        return int(s.value), ph.value, xo.value

    def apply(self, x):
        x = as_cf32(x)
        caps = [int(lib().orc_channelizer_needed_out_count(self.p, i, x.size)) + 1 for i in range(self.n)]
        outs = [np.empty(c, dtype=np.complex64) for c in caps]
        op = (_f32p * self.n)(*[_fptr(o.view(np.float32)) for o in outs])
        counts = (C.c_size_t * self.n)()
        if lib().orc_channelizer_apply(self.p, _fptr(x.view(np.float32)), x.size, op, counts) != 0:
            raise RuntimeError("apply failed")
        return [outs[i][: counts[i]] for i in range(self.n)]


# --------------------------------------------------------------------------
# the reference's own classes (oracle/_ref), when built
# --------------------------------------------------------------------------
_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"{REF_SO} not built (needs the reference sources)")
        L = C.CDLL(REF_SO)
        L.ref_gcd.restype = C.c_uint64
        L.ref_gcd.argtypes = [C.c_uint64, C.c_uint64]
        L.ref_lcm.restype = C.c_uint64
        L.ref_lcm.argtypes = [C.c_uint64, C.c_uint64]
        L.ref_rat_approx.argtypes = [C.c_double, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.ref_freq_shift_params.argtypes = [C.c_uint32, C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
        L.ref_fs_create.restype = _vp
        L.ref_fs_create.argtypes = [C.c_uint32, C.c_int32]
        L.ref_fs_apply.restype = C.c_size_t
        L.ref_fs_apply.argtypes = [_vp, _f32p, C.c_size_t, _f32p]
        L.ref_fs_reset.argtypes = [_vp]
        L.ref_fs_destroy.argtypes = [_vp]
        L.ref_rs_create.restype = _vp
        L.ref_rs_create.argtypes = [C.c_uint32, C.c_uint32, _f32p, C.c_size_t]
        L.ref_rs_needed_out_count.restype = C.c_uint32
        L.ref_rs_needed_out_count.argtypes = [_vp, C.c_uint32]
        L.ref_rs_apply.restype = C.c_int64
        L.ref_rs_apply.argtypes = [_vp, _f32p, C.c_size_t, _f32p, C.c_size_t]
        L.ref_rs_reset.argtypes = [_vp]
        L.ref_rs_destroy.argtypes = [_vp]
        L.ref_fir_create.restype = _vp
        L.ref_fir_create.argtypes = [_f32p, C.c_size_t]
        L.ref_fir_apply.restype = C.c_size_t
        L.ref_fir_apply.argtypes = [_vp, _f32p, C.c_size_t, _f32p]
        L.ref_fir_reset.argtypes = [_vp]
        L.ref_fir_destroy.argtypes = [_vp]
        L.ref_disc_create.restype = _vp
        L.ref_disc_apply.restype = C.c_size_t
        L.ref_disc_apply.argtypes = [_vp, _f32p, C.c_size_t, _f32p]
        L.ref_disc_reset.argtypes = [_vp]
        L.ref_disc_destroy.argtypes = [_vp]
        for name, outp in (("fs", _i16p), ("fir", _i16p), ("disc", _f32p)):
            f = getattr(L, f"ref_{name}_apply_ci16")
            f.restype = C.c_size_t
            f.argtypes = [_vp, _i16p, C.c_size_t, outp]
        L.ref_write_file.restype = C.c_int
        L.ref_write_file.argtypes = [C.c_int, C.c_int, C.c_char_p, _vp, C.c_size_t]
        L.ref_read_file.restype = C.c_int64
        L.ref_read_file.argtypes = [C.c_int, C.c_int, C.c_char_p, _vp, C.c_size_t]
        L.ref_iir_create.restype = _vp
        L.ref_iir_create.argtypes = [_f32p, C.c_size_t, _f32p, C.c_size_t]
        for name in ("ref_iir_apply", "ref_iir_apply_f32"):
            getattr(L, name).restype = C.c_size_t
            getattr(L, name).argtypes = [_vp, _f32p, C.c_size_t, _f32p]
        for name in ("ref_iir_apply_ci16", "ref_iir_apply_i16"):
            getattr(L, name).restype = C.c_size_t
            getattr(L, name).argtypes = [_vp, _i16p, C.c_size_t, _i16p]
        L.ref_iir_reset.argtypes = [_vp]
        L.ref_iir_destroy.argtypes = [_vp]
        L.ref_resample_fir_params.restype = C.c_int64
        L.ref_resample_fir_params.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                              _f32p, C.c_size_t]
        L.ref_resample_fir_params_key.restype = C.c_int64
        L.ref_resample_fir_params_key.argtypes = [C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32,
                                                  C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), _f32p, C.c_size_t]
        L.ref_fm_create.restype = _vp
        L.ref_fm_create.argtypes = [C.c_uint32, C.c_int32, C.c_int]
        L.ref_fm_demod.restype = C.c_int64
        L.ref_fm_demod.argtypes = [_vp, _f32p, C.c_size_t, _i16p, C.c_size_t]
        L.ref_fm_reset.argtypes = [_vp]
        L.ref_fm_destroy.argtypes = [_vp]
        L.ref_chan_create.restype = _vp
        L.ref_chan_create.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_int64), C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(_f32p),
                                      C.POINTER(C.c_uint32)]
        L.ref_chan_apply.restype = C.c_int
        L.ref_chan_apply.argtypes = [_vp, _f32p, C.c_size_t, C.POINTER(_f32p), C.POINTER(C.c_size_t),
                                     C.POINTER(C.c_size_t)]
        L.ref_chan_reset.argtypes = [_vp]
        L.ref_chan_destroy.argtypes = [_vp]
        _ref = L
    return _ref


class RefFreqShift:
    def __init__(self, fs, f):
        self.h = ref().ref_fs_create(fs, f)
        assert self.h

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_fs_destroy(self.h)

    def reset(self):
        ref().ref_fs_reset(self.h)

    def apply(self, x):
        x = as_cf32(x)
        out = np.empty_like(x)
        n = ref().ref_fs_apply(self.h, _fptr(x.view(np.float32)), x.size, _fptr(out.view(np.float32)))
        return out[:n]

    def apply_ci16(self, x):
        x = as_ci16(x)
        out = np.empty_like(x)
        n = ref().ref_fs_apply_ci16(self.h, _iptr(x), x.shape[0], _iptr(out))
        return out[:n]


class RefResampler:
    def __init__(self, up, down, taps):
        t = as_cf32(taps)
        self.h = ref().ref_rs_create(up, down, _fptr(t.view(np.float32)), t.size)
        assert self.h

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_rs_destroy(self.h)

    def reset(self):
        ref().ref_rs_reset(self.h)

    def needed_out_count(self, n):
        return int(ref().ref_rs_needed_out_count(self.h, n))

    def apply(self, x):
        x = as_cf32(x)
        cap = self.needed_out_count(x.size) + 2
        out = np.empty(cap, dtype=np.complex64)
        n = ref().ref_rs_apply(self.h, _fptr(x.view(np.float32)), x.size, _fptr(out.view(np.float32)), cap)
        assert n >= 0
        return out[:n]


class RefFir:
    def __init__(self, coeffs):
        c = as_cf32(coeffs)
        self.h = ref().ref_fir_create(_fptr(c.view(np.float32)), c.size)

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_fir_destroy(self.h)

    def reset(self):
        ref().ref_fir_reset(self.h)

    def apply(self, x):
        x = as_cf32(x)
        out = np.empty_like(x)
        n = ref().ref_fir_apply(self.h, _fptr(x.view(np.float32)), x.size, _fptr(out.view(np.float32)))
        return out[:n]

    def apply_ci16(self, x):
        x = as_ci16(x)
        out = np.empty_like(x)
        n = ref().ref_fir_apply_ci16(self.h, _iptr(x), x.shape[0], _iptr(out))
        return out[:n]


class RefDiscriminator:
    def __init__(self):
        self.h = ref().ref_disc_create()

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_disc_destroy(self.h)

    def reset(self):
        ref().ref_disc_reset(self.h)

    def apply(self, x):
        x = as_cf32(x)
        out = np.empty(x.size, dtype=np.float32)
        n = ref().ref_disc_apply(self.h, _fptr(x.view(np.float32)), x.size, _fptr(out))
        return out[:n]

    def apply_ci16(self, x):
        x = as_ci16(x)
        out = np.empty(x.shape[0], dtype=np.float32)
        n = ref().ref_disc_apply_ci16(self.h, _iptr(x), x.shape[0], _fptr(out))
        return out[:n]


class RefIir:
    def __init__(self, b, a):
        self._b, self._a = as_cf32(b), as_cf32(a)
        self.h = ref().ref_iir_create(_fptr(self._b.view(np.float32)), self._b.size,
                                      _fptr(self._a.view(np.float32)), self._a.size)
        assert self.h

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_iir_destroy(self.h)

    def reset(self):
        ref().ref_iir_reset(self.h)

    def apply(self, x):
        x = as_cf32(x)
        out = np.empty_like(x)
        n = ref().ref_iir_apply(self.h, _fptr(x.view(np.float32)), x.size, _fptr(out.view(np.float32)))
        return out[:n]

    def apply_f32(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty_like(x)
        n = ref().ref_iir_apply_f32(self.h, _fptr(x), x.size, _fptr(out))
        return out[:n]

    def apply_ci16(self, x):
        x = as_ci16(x)
        out = np.empty_like(x)
        n = ref().ref_iir_apply_ci16(self.h, _iptr(x), x.shape[0], _iptr(out))
        return out[:n]

    def apply_i16(self, x):
        x = np.ascontiguousarray(x, dtype=np.int16)
        out = np.empty_like(x)
        n = ref().ref_iir_apply_i16(self.h, _iptr(x), x.size, _iptr(out))
        return out[:n]


def ref_predefined_filter(in_rate, out_rate, taps, source):
    """The reference's predefined-table entry (taps: filter_taps_e value, source:
    0 REMEZ, 1 FIRLS) -> (up, down, complex64 taps), or None when absent."""
    up, down = C.c_uint32(), C.c_uint32()
    c = np.empty(4096, dtype=np.complex64)
    n = ref().ref_resample_fir_params_key(in_rate, out_rate, taps, source, C.byref(up), C.byref(down),
                                          _fptr(c.view(np.float32)), c.size)
    if n < 0:
        return None
    return up.value, down.value, c[:n].copy()


def ref_resample_fir_params(in_rate, out_rate):
    """The reference's predefined 64-tap REMEZ resampling filter -> (up, down, taps)."""
    up, down = C.c_uint32(), C.c_uint32()
    taps = np.empty(4096, dtype=np.complex64)
    n = ref().ref_resample_fir_params(in_rate, out_rate, C.byref(up), C.byref(down),
                                      _fptr(taps.view(np.float32)), taps.size)
    if n < 0:
        raise KeyError((in_rate, out_rate))
    return up.value, down.value, taps[:n].copy()


class RefFm:
    def __init__(self, fs, offset_hz, europe=False):
        self.h = ref().ref_fm_create(fs, offset_hz, 1 if europe else 0)
        assert self.h

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_fm_destroy(self.h)

    def reset(self):
        ref().ref_fm_reset(self.h)

    def demod(self, x):
        x = as_cf32(x)
        cap = x.size
        out = np.empty(max(cap, 1), dtype=np.int16)
        n = ref().ref_fm_demod(self.h, _fptr(x.view(np.float32)), x.size, _iptr(out), cap)
        return None if n < 0 else out[:n]


class RefChannelizer:
    def __init__(self, fs, channels):
        n = len(channels)
        self.n = n
        self._taps = [as_cf32(c["taps"]) for c in channels]
        shifts = (C.c_int64 * n)(*[int(c["freq_shift_in_hz"]) for c in channels])
        rates = (C.c_uint32 * n)(*[int(c.get("output_sample_rate_in_sps", 0)) for c in channels])
        ups = (C.c_uint32 * n)(*[int(c["up_rate"]) for c in channels])
        downs = (C.c_uint32 * n)(*[int(c["down_rate"]) for c in channels])
        tp = (_f32p * n)(*[_fptr(t.view(np.float32)) for t in self._taps])
        nt = (C.c_uint32 * n)(*[t.size for t in self._taps])
        self.h = ref().ref_chan_create(fs, n, shifts, rates, ups, downs, tp, nt)
        assert self.h
        self._ups = [int(c["up_rate"]) for c in channels]
        self._downs = [int(c["down_rate"]) for c in channels]

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_chan_destroy(self.h)

    def reset(self):
        ref().ref_chan_reset(self.h)

    def apply(self, x):
        x = as_cf32(x)
        caps = [x.size * u // d + 2 for u, d in zip(self._ups, self._downs)]
        outs = [np.empty(c, dtype=np.complex64) for c in caps]
        op = (_f32p * self.n)(*[_fptr(o.view(np.float32)) for o in outs])
        cp = (C.c_size_t * self.n)(*caps)
        counts = (C.c_size_t * self.n)()
        if ref().ref_chan_apply(self.h, _fptr(x.view(np.float32)), x.size, op, cp, counts) != 0:
            raise RuntimeError("apply failed")
        return [outs[i][: counts[i]] for i in range(self.n)]


# --------------------------------------------------------------------------
# FP64 closed forms (independent of both C libraries)
# --------------------------------------------------------------------------
def fp64_freq_shift(x, fs, f, start_index=0, rollover=None):
    """x[n]*exp(j*w*((start+n) mod R)) in complex128; R from the reference's rule
    unless given. With rollover=0 the phase runs continuously (ideal NCO)."""
    x = np.asarray(x, dtype=np.complex128)
    if f == 0:
        return x.copy()
    R, w = freq_shift_params(fs, f) if rollover is None else (rollover, 2 * math.pi * f / fs)
    n = np.arange(x.size, dtype=np.int64) + int(start_index)
    if R:
        n = n % R
    # exact reduction of f*n/fs in integers keeps the FP64 phase exact
    # (n < 2^32 and |f| < 2^31, so n*f fits in int64)
    ph = (n * int(f)) % int(fs)
    return x * np.exp(2j * np.pi * ph.astype(np.float64) / fs)


def fp64_resample(x, taps, up, down, n_out=None):
    """Brute force: zero-stuff by up, full convolution, keep every down-th
    sample; the first n_out outputs (the reference's streaming count)."""
    x = np.asarray(x, dtype=np.complex128)
    h = np.real(np.asarray(taps, dtype=np.complex128))
    z = np.zeros(x.size * up, dtype=np.complex128)
    z[::up] = x
    y = np.convolve(z, h)[::down]
    if n_out is None:
        n_out = -(-x.size * up // down)
    return y[:n_out]


def fp64_fir(x, h):
    x = np.asarray(x, dtype=np.complex128)
    return np.convolve(x, np.asarray(h, dtype=np.complex128))[: x.size]


def fp64_discriminator(x, prev=None):
    x = np.asarray(x, dtype=np.complex128)
    if prev is not None:
        x = np.concatenate([[prev], x])
    return np.angle(x[1:] * np.conj(x[:-1]))
