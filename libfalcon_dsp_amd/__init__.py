"""libfalcon_dsp_amd -- MI355X (gfx950) streaming channelizer.

The NCO frequency shift, polyphase resampler, FIR and IIR filters, polar
discriminator, multi-rate channelizer and FM mono demodulator of
libfalcon_dsp, computed by hand-written HIP kernels behind the C ABI in
include/fdsp.h.
"""
from ._native import FdspError, LIB_PATH, lib  # noqa: F401
from .dsp import (  # noqa: F401
    CI16, DeviceBuffer, Event, FirBank, FmDemodulator, Iir1, IirBank, IirFilter, fm_deemphasis, iir_filter, resample_fir_params, FILTER_REMEZ, FILTER_FIRLS, FILTER_TAPS_OPTIMAL, FirFilter, FreqShift, MultiRateChannelizer, PolarDiscriminator,
    PolyphaseResampler, Server, Stream, TILING_CLAIM, TILING_STRIDED, calculate_gcd, calculate_lcm, device_count,
    device_sync, fir_filter, freq_shift, get_freq_shift_params, polar_discriminator, rat_approx, reduce_rates, resample, upfirdn,
)

__version__ = "0.1.0"
