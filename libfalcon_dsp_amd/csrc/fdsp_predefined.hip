// Predefined anti-alias filters for resampling: the lookup behind
// get_resample_fir_params (reference src/transform/falcon_dsp_fir_filter.cc:
// 111-130). The table is generated at build time by
// libfalcon_dsp_amd/predefined.py, which designs each filter by the
// reference's documented procedure. Host code only.
#include <cstdint>
#include <cstring>

#include "../../include/fdsp.h"
#include "fdsp_host.h"

using namespace fdsp;

namespace {

struct PredefinedFilter {
    uint32_t in_rate, out_rate;
    int32_t taps;     // FDSP_FILTER_TAPS_* value
    uint32_t source;  // FDSP_FILTER_REMEZ or FDSP_FILTER_FIRLS
    uint32_t up, down, n;
    const float *coeffs;  // real taps (the imaginary parts are zero)
};

#include "fdsp_predefined_table.inc"

}  // namespace

extern "C" int fdsp_resample_fir_params(uint32_t in_rate, uint32_t out_rate, int32_t taps, uint32_t source,
                                        uint32_t *up, uint32_t *down, fdsp_cf32 *coeffs, size_t cap,
                                        size_t *n_coeffs)
{
    if (!up || !down || !n_coeffs)
        return set_err(FDSP_EINVAL, "resample_fir_params: null output");
    *n_coeffs = 0;
    for (const PredefinedFilter &f : kPredefined) {
        if (f.in_rate != in_rate || f.out_rate != out_rate || f.taps != taps || f.source != source)
            continue;
        *up = f.up;
        *down = f.down;
        *n_coeffs = f.n;
        if (!coeffs || cap < f.n)
            return set_err(FDSP_ERANGE, "resample_fir_params: %u coefficients, room for %zu", f.n, cap);
        for (uint32_t i = 0; i < f.n; ++i)
            coeffs[i] = fdsp_cf32{f.coeffs[i], 0.f};
        return FDSP_OK;
    }
    return set_err(FDSP_EINVAL, "resample_fir_params: no predefined filter for %u -> %u sps, taps %d, source %u",
                   in_rate, out_rate, taps, source);
}
