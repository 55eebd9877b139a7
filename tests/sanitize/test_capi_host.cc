// Host-layer run of the C ABI and the drop-in utilities under AddressSanitizer
// and UndefinedBehaviorSanitizer, built by tests/test_sanitizers_cpu.py from
// the library's sources (host code only; no GPU is needed or used).
//
// It drives every path that needs no device: the pure host functions (gcd,
// lcm, rational approximation, NCO parameters, the predefined filter table,
// FM de-emphasis, the drop-in file readers and writers), argument validation
// and null-handle errors of every handle type, the error message formatting,
// and, where no device is present, the failure and clean-up paths of
// constructors whose first device call fails.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "fdsp.h"
#include "utilities/falcon_dsp_utils.h"

static int g_fail = 0;
#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::fprintf(stderr, "CHECK failed: %s (line %d)\n", #c, __LINE__); \
            g_fail = 1;                                                      \
        }                                                                    \
    } while (0)

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    CHECK(fdsp_version() > 0);
    CHECK(fdsp_gcd(48, 18) == 6 && fdsp_lcm(4, 6) == 12 && fdsp_gcd(0, 7) == 7);
    int64_t num = 0, den = 0;
    fdsp_rat_approx(0.75, 1024, &num, &den);
    CHECK(num == 3 && den == 4);
    fdsp_rat_approx(M_PI, 1000, &num, &den);
    CHECK(den > 0 && den <= 1000 && std::fabs((double)num / den - M_PI) < 1e-5);
    uint32_t roll = 0;
    double omega = 0;
    // the period fs/f = 43.48 samples, scaled by ten until at least 1e5
    CHECK(fdsp_freq_shift_params(1000000, 23000, &roll, &omega) == FDSP_OK && roll == 434783);
    CHECK(fdsp_freq_shift_params(1000000, 0, &roll, &omega) == FDSP_OK && roll == 1 && omega == 0.0);
    CHECK(fdsp_freq_shift_params(0, 5, &roll, &omega) != FDSP_OK);
    float b0 = 0, a1 = 0;
    CHECK(fdsp_fm_deemphasis(75e-6, &b0, &a1) == FDSP_OK && b0 > 0 && a1 < 0);
    CHECK(fdsp_fm_deemphasis(-1.0, &b0, &a1) == FDSP_EINVAL);
    CHECK(fdsp_last_error() && std::string(fdsp_last_error()).size() > 0);

    // predefined filters: sizes, a too-small buffer, an unknown pair
    for (int32_t taps : {32, 64, 128, 256, FDSP_FILTER_TAPS_OPTIMAL})
        for (uint32_t src : {FDSP_FILTER_REMEZ, FDSP_FILTER_FIRLS}) {
            uint32_t up = 0, down = 0;
            size_t n = 0;
            std::vector<fdsp_cf32> c(1024);
            const int rc = fdsp_resample_fir_params(1140000, 228000, taps, src, &up, &down, c.data(), c.size(), &n);
            CHECK(rc == FDSP_OK || rc == FDSP_EINVAL);
            if (rc == FDSP_OK) {
                CHECK(n > 0 && n <= c.size() && up == 1 && down == 5);
                size_t n2 = 0;
                CHECK(fdsp_resample_fir_params(1140000, 228000, taps, src, &up, &down, c.data(), 3, &n2) ==
                      FDSP_ERANGE);
                CHECK(n2 == n);
            }
        }
    {
        uint32_t up, down;
        size_t n;
        fdsp_cf32 c[4];
        CHECK(fdsp_resample_fir_params(1234, 567, 64, FDSP_FILTER_REMEZ, &up, &down, c, 4, &n) == FDSP_EINVAL);
    }

    // null handles and bad arguments, for every handle type
    CHECK(fdsp_freq_shift_reset(nullptr) == FDSP_ENOTINIT);
    CHECK(fdsp_freq_shift_create(1000, nullptr, 1, nullptr) == FDSP_EINVAL);
    CHECK(fdsp_resampler_reset(nullptr) == FDSP_ENOTINIT);
    fdsp_resampler_t r = nullptr;
    const fdsp_cf32 taps4[4] = {{1, 0}, {0.5f, 0}, {0.25f, 0}, {0.125f, 0}};
    CHECK(fdsp_resampler_create(0, 1, taps4, 4, &r) == FDSP_EINVAL && r == nullptr);
    CHECK(fdsp_resampler_set_state(nullptr, 0, 0, nullptr, 0) == FDSP_ENOTINIT);
    CHECK(fdsp_fir_create(nullptr, 0, nullptr) == FDSP_EINVAL);
    CHECK(fdsp_fir_bank_reset(nullptr) == FDSP_ENOTINIT);
    CHECK(fdsp_discriminator_reset(nullptr) == FDSP_ENOTINIT);
    fdsp_iir_t iir = nullptr;
    const fdsp_cf32 one[1] = {{1, 0}}, zero_a[2] = {{0, 0}, {1, 0}};
    CHECK(fdsp_iir_create(one, 1, zero_a, 2, &iir) == FDSP_EINVAL && iir == nullptr);
    CHECK(fdsp_iir_create(one, 300, one, 1, &iir) == FDSP_EINVAL);
    CHECK(fdsp_iir_reset(nullptr) == FDSP_ENOTINIT);
    CHECK(fdsp_iir_apply_device_noalloc(nullptr, nullptr, 1, nullptr, nullptr) == FDSP_ENOTINIT);
    fdsp_iir_bank_t bank = nullptr;
    CHECK(fdsp_iir_bank_create(one, 1, one, 1, 0, &bank) == FDSP_EINVAL && bank == nullptr);
    CHECK(fdsp_iir_bank_create(one, 18, one, 1, 4, &bank) == FDSP_EINVAL);
    CHECK(fdsp_iir_bank_apply_device(nullptr, nullptr, 1, 1, nullptr, 1, nullptr) == FDSP_ENOTINIT);
    CHECK(fdsp_iir1_reset(nullptr) == FDSP_ENOTINIT);
    CHECK(fdsp_fm_demod_reset(nullptr) == FDSP_ENOTINIT);
    fdsp_channelizer_t ch = nullptr;
    CHECK(fdsp_channelizer_create(0, nullptr, 0, FDSP_CHAN_BATCHED, &ch) == FDSP_EINVAL && ch == nullptr);
    fdsp_channel_desc bad{};
    bad.up_rate = 1;
    bad.down_rate = 0;
    bad.taps = taps4;
    bad.n_taps = 4;
    CHECK(fdsp_channelizer_create(1000000, &bad, 1, FDSP_CHAN_BATCHED, &ch) == FDSP_EINVAL);
    CHECK(fdsp_channelizer_create(1000000, &bad, 1, 7, &ch) == FDSP_EINVAL);
    CHECK(fdsp_channelizer_reset(nullptr) == FDSP_ENOTINIT);
    CHECK(fdsp_channelizer_autotune(nullptr, nullptr, 0, 1, nullptr, 0, nullptr) == FDSP_ENOTINIT);
    fdsp_serve_config cfg{};
    fdsp_serve_t srv = nullptr;
    CHECK(fdsp_serve_create(&cfg, &srv) == FDSP_EINVAL && srv == nullptr);
    cfg.fs = 1000;
    cfg.max_streams = 4;
    cfg.max_chunk = 100;
    cfg.max_batch = 1u << 31;
    CHECK(fdsp_serve_create(&cfg, &srv) == FDSP_ERANGE);
    CHECK(fdsp_serve_poll(nullptr, 1, 0, nullptr) == FDSP_ENOTINIT);
    CHECK(fdsp_serve_destroy(nullptr) == FDSP_OK);

    // constructors on a machine without a device fail cleanly; with one they succeed
    int devices = 0;
    const bool have_gpu = fdsp_device_count(&devices) == FDSP_OK && devices > 0;
    if (!have_gpu) {
        fdsp_channel_desc d{};
        d.output_sample_rate_in_sps = 250000;
        d.freq_shift_in_hz = 1000;
        d.up_rate = 1;
        d.down_rate = 4;
        d.taps = taps4;
        d.n_taps = 4;
        CHECK(fdsp_channelizer_create(1000000, &d, 1, FDSP_CHAN_BATCHED, &ch) != FDSP_OK && ch == nullptr);
        CHECK(fdsp_iir_create(one, 1, one, 1, &iir) != FDSP_OK && iir == nullptr);
        CHECK(fdsp_iir_bank_create(one, 1, one, 1, 8, &bank) != FDSP_OK && bank == nullptr);
        CHECK(fdsp_resampler_create(1, 2, taps4, 4, &r) != FDSP_OK && r == nullptr);
        cfg.max_batch = 1000;
        CHECK(fdsp_serve_create(&cfg, &srv) != FDSP_OK && srv == nullptr);
    }

    // drop-in file utilities: every format round trip
    using namespace falcon_dsp;
    std::vector<std::complex<float>> cf{{1.5f, -2.25f}, {0.f, 3.f}, {-7.f, 0.125f}};
    std::vector<std::complex<int16_t>> ci{{1, -2}, {-32768, 32767}, {0x80, -0x81}};
    std::vector<float> fl{0.5f, -1.f, 3.25f};
    std::vector<int16_t> i16{1, -300, 0x7f80, -2};
    for (file_type_e t : {file_type_e::BINARY, file_type_e::ASCII}) {
        const std::string f = dir + "/fdsp_sanitize_" + std::to_string((int)t);
        std::vector<std::complex<int16_t>> ci2;
        CHECK(write_complex_data_to_file(f + "_ci", t, ci) && read_complex_data_from_file(f + "_ci", t, ci2));
        CHECK(ci2 == ci);
        std::vector<int16_t> i2;
        CHECK(write_data_to_file(f + "_i", t, i16) && read_data_from_file(f + "_i", t, i2));
        CHECK(i2 == i16);
        std::vector<std::complex<float>> cf2;
        CHECK(write_complex_data_to_file(f + "_cf", t, cf));
        CHECK(read_complex_data_from_file(f + "_cf", t, cf2) && cf2 == cf);
        std::vector<float> fl2;
        CHECK(write_data_to_file(f + "_f", t, fl));
        const bool okf = read_data_from_file(f + "_f", t, fl2);
        CHECK(t == file_type_e::BINARY ? !okf : (okf && fl2.size() == fl.size()));
        for (const char *suf : {"_ci", "_i", "_cf", "_f"})
            std::remove((f + suf).c_str());
    }
    std::vector<int16_t> none;
    CHECK(!read_data_from_file(dir + "/fdsp_sanitize_missing_file", file_type_e::BINARY, none));
    uint64_t g = calculate_gcd(1140000, 228000);
    CHECK(g == 228000);
    int64_t n2 = 0, d2 = 0;
    rat_approx(228000.0 / 1140000.0, 1024, n2, d2);
    CHECK(n2 == 1 && d2 == 5);

    if (g_fail) {
        std::fprintf(stderr, "host-layer checks failed\n");
        return 1;
    }
    std::printf("host-layer checks passed (device present: %d)\n", have_gpu ? 1 : 0);
    return 0;
}
