"""ctypes binding of libfdsp_hip.so (the C ABI in include/fdsp.h).

The shared library is built in-tree by ``libfalcon_dsp_amd.build``. There is
no fallback: if the library is missing or fails to load, every entry point
raises, so a run can never silently compute on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfdsp_hip.so")

FDSP_OK = 0
FDSP_EINVAL = -1
FDSP_ENOTINIT = -2
FDSP_EBUSY = -3
FDSP_ENOMEM = -4
FDSP_EHIP = -5
FDSP_ERANGE = -6
FDSP_EAGAIN = -7

FDSP_CHAN_PLAIN = 0
FDSP_CHAN_BATCHED = 1
FDSP_CHAN_BATCHED_BF16X2 = 2
FDSP_KERNEL_SEGMENT_F32 = 0
FDSP_KERNEL_GROUPS_F32 = 1
FDSP_KERNEL_GROUPS_BF16X2 = 2
FDSP_KERNEL_PLAIN = 3
KERNEL_NAMES = {0: "chan_seg_kernel", 1: "chan_groups_kernel", 2: "chan_groups_kernel", 3: "nco_multi_kernel+resampler_kernel"}
FDSP_TAIL_NONE = 0
FDSP_TAIL_DISCRIMINATOR = 1

_NAMES = {
    FDSP_EINVAL: "FDSP_EINVAL", FDSP_ENOTINIT: "FDSP_ENOTINIT", FDSP_EBUSY: "FDSP_EBUSY",
    FDSP_ENOMEM: "FDSP_ENOMEM", FDSP_EHIP: "FDSP_EHIP", FDSP_ERANGE: "FDSP_ERANGE",
    FDSP_EAGAIN: "FDSP_EAGAIN",
}


class FdspError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_NAMES.get(code, code)}: {msg}")
        self.code = code


class ChannelDesc(C.Structure):
    """fdsp_channel_desc (mirrors multi_rate_channelizer_channel_s)."""
    _fields_ = [
        ("output_sample_rate_in_sps", C.c_uint32),
        ("freq_shift_in_hz", C.c_int64),
        ("up_rate", C.c_uint32),
        ("down_rate", C.c_uint32),
        ("taps", C.c_void_p),
        ("n_taps", C.c_uint32),
        ("tail", C.c_uint32),
    ]


class ServeConfig(C.Structure):
    """fdsp_serve_config."""
    _fields_ = [
        ("fs", C.c_uint32),
        ("max_streams", C.c_uint32),
        ("max_chunk", C.c_uint32),
        ("max_batch", C.c_uint32),
        ("hbm_budget", C.c_uint64),
        ("state_bytes", C.c_uint64),
    ]


class ServeResult(C.Structure):
    """fdsp_serve_result."""
    _fields_ = [
        ("n_out", C.c_size_t),
        ("status", C.c_int),
        ("submit_ns", C.c_uint64),
        ("done_ns", C.c_uint64),
        ("batch", C.c_uint64),
        ("pack_ns", C.c_uint64),
        ("issue_ns", C.c_uint64),
        ("gpu_ns", C.c_uint64),
    ]


class ServeStats(C.Structure):
    """fdsp_serve_stats."""
    _fields_ = [
        ("hbm_used", C.c_uint64),
        ("hbm_budget", C.c_uint64),
        ("live_streams", C.c_uint32),
        ("queued", C.c_size_t),
        ("batches", C.c_uint64),
        ("chunks", C.c_uint64),
        ("in_samples", C.c_uint64),
        ("out_samples", C.c_uint64),
        ("state_pool", C.c_uint64),
    ]


class TilingTrial(C.Structure):
    """fdsp_tiling_trial."""
    _fields_ = [("segment_len", C.c_uint32), ("block_threads", C.c_uint32), ("us", C.c_float),
                ("flags", C.c_uint32)]


_vp = C.c_void_p
_u32 = C.c_uint32
_u64 = C.c_uint64
_sz = C.c_size_t
_i = C.c_int
_P = C.POINTER

# (name, restype, argtypes)
_SIGS = [
    ("fdsp_last_error", C.c_char_p, []),
    ("fdsp_version", _i, []),
    ("fdsp_device_count", _i, [_P(_i)]),
    ("fdsp_set_device", _i, [_i]),
    ("fdsp_malloc", _i, [_P(_vp), _sz]),
    ("fdsp_free", _i, [_vp]),
    ("fdsp_memcpy_h2d", _i, [_vp, _vp, _sz, _vp]),
    ("fdsp_memcpy_d2h", _i, [_vp, _vp, _sz, _vp]),
    ("fdsp_memcpy_d2d", _i, [_vp, _vp, _sz, _vp]),
    ("fdsp_memset", _i, [_vp, _i, _sz, _vp]),
    ("fdsp_host_alloc_pinned", _i, [_P(_vp), _sz]),
    ("fdsp_host_free_pinned", _i, [_vp]),
    ("fdsp_stream_create", _i, [_P(_vp)]),
    ("fdsp_stream_destroy", _i, [_vp]),
    ("fdsp_stream_sync", _i, [_vp]),
    ("fdsp_device_sync", _i, []),
    ("fdsp_event_create", _i, [_P(_vp)]),
    ("fdsp_event_destroy", _i, [_vp]),
    ("fdsp_event_record", _i, [_vp, _vp]),
    ("fdsp_event_sync", _i, [_vp]),
    ("fdsp_event_elapsed_ms", _i, [_vp, _vp, _P(C.c_float)]),
    ("fdsp_gcd", _u64, [_u64, _u64]),
    ("fdsp_lcm", _u64, [_u64, _u64]),
    ("fdsp_rat_approx", None, [C.c_double, C.c_int64, _P(C.c_int64), _P(C.c_int64)]),
    ("fdsp_freq_shift_params", _i, [_u32, C.c_int32, _P(_u32), _P(C.c_double)]),
    ("fdsp_freq_shift_create", _i, [_u32, _P(C.c_int32), _u32, _P(_vp)]),
    ("fdsp_freq_shift_destroy", _i, [_vp]),
    ("fdsp_freq_shift_reset", _i, [_vp]),
    ("fdsp_freq_shift_get_state", _i, [_vp, _u32, _P(_u64)]),
    ("fdsp_freq_shift_apply", _i, [_vp, _vp, _sz, _P(_vp)]),
    ("fdsp_freq_shift_apply_device", _i, [_vp, _vp, _sz, _vp, _sz, _vp]),
    ("fdsp_freq_shift_apply_ci16", _i, [_vp, _vp, _sz, _P(_vp)]),
    ("fdsp_freq_shift_apply_ci16_device", _i, [_vp, _vp, _sz, _vp, _sz, _vp]),
    ("fdsp_resampler_create", _i, [_u32, _u32, _vp, _u32, _P(_vp)]),
    ("fdsp_resampler_destroy", _i, [_vp]),
    ("fdsp_resampler_reset", _i, [_vp]),
    ("fdsp_resampler_needed_out_count", _i, [_vp, _u32, _P(_u32)]),
    ("fdsp_resampler_coeffs_per_phase", _i, [_vp, _P(_u32)]),
    ("fdsp_resampler_get_state", _i, [_vp, _P(_u32), _P(_u32), _P(_u64), _P(_u64)]),
    ("fdsp_resampler_apply", _i, [_vp, _vp, _sz, _vp, _sz, _P(_sz)]),
    ("fdsp_resampler_apply_device", _i, [_vp, _vp, _sz, _vp, _sz, _P(_sz), _vp]),
    ("fdsp_fir_create", _i, [_vp, _u32, _P(_vp)]),
    ("fdsp_fir_destroy", _i, [_vp]),
    ("fdsp_fir_reset", _i, [_vp]),
    ("fdsp_fir_apply", _i, [_vp, _vp, _sz, _vp]),
    ("fdsp_fir_apply_device", _i, [_vp, _vp, _sz, _vp, _vp]),
    ("fdsp_fir_apply_ci16", _i, [_vp, _vp, _sz, _vp]),
    ("fdsp_fir_apply_ci16_device", _i, [_vp, _vp, _sz, _vp, _vp]),
    ("fdsp_fir_bank_create", _i, [_u32, _u32, _vp, _P(_vp)]),
    ("fdsp_fir_bank_destroy", _i, [_vp]),
    ("fdsp_fir_bank_reset", _i, [_vp]),
    ("fdsp_fir_bank_apply_device", _i, [_vp, _vp, _sz, _sz, _vp, _sz, _vp]),
    ("fdsp_discriminator_create", _i, [_P(_vp)]),
    ("fdsp_discriminator_destroy", _i, [_vp]),
    ("fdsp_discriminator_reset", _i, [_vp]),
    ("fdsp_discriminator_apply", _i, [_vp, _vp, _sz, _vp, _P(_sz)]),
    ("fdsp_discriminator_apply_device", _i, [_vp, _vp, _sz, _vp, _P(_sz), _vp]),
    ("fdsp_discriminator_apply_ci16", _i, [_vp, _vp, _sz, _vp, _P(_sz)]),
    ("fdsp_discriminator_apply_ci16_device", _i, [_vp, _vp, _sz, _vp, _P(_sz), _vp]),
    ("fdsp_iir_create", _i, [_vp, _u32, _vp, _u32, _P(_vp)]),
    ("fdsp_iir_destroy", _i, [_vp]),
    ("fdsp_iir_reset", _i, [_vp]),
    ("fdsp_iir_info", _i, [_vp, _P(_u32), _P(_u32), _P(_u32)]),
    ("fdsp_iir_set_mode", _i, [_vp, C.c_int]),
    ("fdsp_iir_apply", _i, [_vp, _vp, _sz, _vp]),
    ("fdsp_iir_apply_f32", _i, [_vp, _vp, _sz, _vp]),
    ("fdsp_iir_apply_ci16", _i, [_vp, _vp, _sz, _vp]),
    ("fdsp_iir_apply_i16", _i, [_vp, _vp, _sz, _vp]),
    ("fdsp_iir_apply_device", _i, [_vp, _vp, _sz, _vp, _vp]),
    ("fdsp_iir_apply_device_noalloc", _i, [_vp, _vp, _sz, _vp, _vp]),
    ("fdsp_iir_apply_f32_device", _i, [_vp, _vp, _sz, _vp, _vp]),
    ("fdsp_iir_apply_ci16_device", _i, [_vp, _vp, _sz, _vp, _vp]),
    ("fdsp_iir_apply_i16_device", _i, [_vp, _vp, _sz, _vp, _vp]),
    ("fdsp_iir_bank_create", _i, [_vp, _u32, _vp, _u32, _u32, _P(_vp)]),
    ("fdsp_iir_bank_destroy", _i, [_vp]),
    ("fdsp_iir_bank_reset", _i, [_vp]),
    ("fdsp_iir_bank_apply", _i, [_vp, _vp, _sz, _vp]),
    ("fdsp_iir_bank_apply_ci16", _i, [_vp, _vp, _sz, _vp]),
    ("fdsp_iir_bank_apply_device", _i, [_vp, _vp, _sz, _sz, _vp, _sz, _vp]),
    ("fdsp_iir_bank_apply_ci16_device", _i, [_vp, _vp, _sz, _sz, _vp, _sz, _vp]),
    ("fdsp_resample_fir_params", _i, [_u32, _u32, C.c_int32, _u32, _P(_u32), _P(_u32), _vp, _sz, _P(_sz)]),
    ("fdsp_iir1_create", _i, [C.c_float, C.c_float, _P(_vp)]),
    ("fdsp_iir1_destroy", _i, [_vp]),
    ("fdsp_iir1_reset", _i, [_vp]),
    ("fdsp_iir1_apply", _i, [_vp, _vp, _sz, _vp]),
    ("fdsp_iir1_apply_device", _i, [_vp, _vp, _sz, _vp, _vp]),
    ("fdsp_fm_deemphasis", _i, [C.c_double, _P(C.c_float), _P(C.c_float)]),
    ("fdsp_fm_demod_create", _i, [_u32, C.c_int32, C.c_double, _u32, _u32, _vp, _u32, _u32, _u32, _vp, _u32,
                                  _P(_vp)]),
    ("fdsp_fm_demod_destroy", _i, [_vp]),
    ("fdsp_fm_demod_reset", _i, [_vp]),
    ("fdsp_fm_demod_out_count", _i, [_vp, _sz, _P(_sz)]),
    ("fdsp_fm_demod_apply", _i, [_vp, _vp, _sz, _vp, _vp, _sz, _P(_sz)]),
    ("fdsp_fm_demod_apply_device", _i, [_vp, _vp, _sz, _vp, _vp, _sz, _P(_sz), _vp]),
    ("fdsp_channelizer_create", _i, [_u32, _P(ChannelDesc), _u32, _i, _P(_vp)]),
    ("fdsp_channelizer_destroy", _i, [_vp]),
    ("fdsp_channelizer_reset", _i, [_vp]),
    ("fdsp_channelizer_num_channels", _i, [_vp, _P(_u32)]),
    ("fdsp_channelizer_needed_out_count", _i, [_vp, _u32, _u32, _P(_u32)]),
    ("fdsp_channelizer_get_state", _i, [_vp, _u32, _P(_u64), _P(_u32), _P(_u32), _P(_u64), _P(_u64)]),
    ("fdsp_channelizer_apply", _i, [_vp, _vp, _sz, _P(_vp), _P(_sz), _P(_sz)]),
    ("fdsp_channelizer_apply_device", _i, [_vp, _vp, _sz, _P(_vp), _P(_sz), _P(_sz), _vp]),
    ("fdsp_channelizer_graph_create", _i, [_vp, _sz, _vp, _P(_vp), _P(_sz)]),
    ("fdsp_channelizer_graph_launch", _i, [_vp, _vp]),
    ("fdsp_channelizer_graph_counts", _i, [_vp, _P(_sz)]),
    ("fdsp_channelizer_graph_destroy", _i, [_vp]),
    ("fdsp_channelizer_set_tiling", _i, [_vp, _u32, _u32]),
    ("fdsp_channelizer_get_tiling", _i, [_vp, _P(_u32), _P(_u32)]),
    ("fdsp_freq_shift_apply_chunks_device", _i, [_vp, _vp, _sz, _sz, _vp, _sz, _vp]),
    ("fdsp_freq_shift_set_state", _i, [_vp, _u64]),
    ("fdsp_resampler_apply_chunks_device", _i, [_vp, _vp, _sz, _sz, _vp, _sz, _P(_sz), _vp]),
    ("fdsp_resampler_set_state", _i, [_vp, _u32, _u32, _vp, _sz]),
    ("fdsp_channelizer_kernel_info", _i, [_vp, _P(_u32), _P(_u32), _P(_u32)]),
    ("fdsp_channelizer_set_kernel", _i, [_vp, _u32]),
    ("fdsp_channelizer_plan_info", _i, [_vp, _sz, _P(_u32), _P(_u32), _P(_u32), _P(_u32)]),
    ("fdsp_channelizer_autotune", _i, [_vp, _vp, _sz, _u32, _P(TilingTrial), _sz, _P(_sz)]),
    ("fdsp_channelizer_set_fusion", _i, [_vp, _i]),
    ("fdsp_channelizer_get_fusion", _i, [_vp, _P(_i), _P(_u32)]),
    ("fdsp_serve_create", _i, [_P(ServeConfig), _P(_vp)]),
    ("fdsp_serve_destroy", _i, [_vp]),
    ("fdsp_serve_admit", _i, [_vp, _P(ChannelDesc), _P(_u32)]),
    ("fdsp_serve_max_out", _i, [_vp, _u32, _sz, _P(_sz)]),
    ("fdsp_serve_submit", _i, [_vp, _u32, _vp, _sz, _vp, _sz, _P(_u64)]),
    ("fdsp_serve_poll", _i, [_vp, _u64, _i, _P(ServeResult)]),
    ("fdsp_serve_release", _i, [_vp, _u32]),
    ("fdsp_serve_info", _i, [_vp, _P(ServeStats)]),
]

SYMBOLS = [s[0] for s in _SIGS]

_lib = None


def lib():
    """Load libfdsp_hip.so (raises if it is absent: there is no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().fdsp_last_error().decode(errors="replace")


def check(rc: int):
    if rc != FDSP_OK:
        raise FdspError(rc, last_error())
    return rc
