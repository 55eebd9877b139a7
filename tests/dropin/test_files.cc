// test_files.cc -- the drop-in's sample-file functions (reference
// src/utilities/falcon_dsp_utils.cc:301-811) as a converter, for the CPU
// tests: read a file with one reader, write it with one writer. No GPU use.
//   test_files <ci16|cf32|i16|f32> <bin|ascii> <in> <bin|ascii> <out>
// Exit status 0 when both calls return true, 2 when the read fails, 3 when
// the write fails, 1 on bad arguments.
#include <complex>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "utilities/falcon_dsp_utils.h"

using namespace falcon_dsp;

static bool fmt(const char *s, file_type_e *t)
{
    if (!std::strcmp(s, "bin"))
        *t = file_type_e::BINARY;
    else if (!std::strcmp(s, "ascii"))
        *t = file_type_e::ASCII;
    else
        return false;
    return true;
}

template <typename V, typename R, typename W>
static int convert(file_type_e ti, const char *in, file_type_e to, const char *out, R rd, W wr)
{
    V v;
    if (!rd(in, ti, v))
        return 2;
    return wr(out, to, v) ? 0 : 3;
}

int main(int argc, char **argv)
{
    file_type_e ti, to;
    if (argc != 6 || !fmt(argv[2], &ti) || !fmt(argv[4], &to))
        return 1;
    const std::string kind = argv[1];
    if (kind == "ci16")
        return convert<std::vector<std::complex<int16_t>>>(
            ti, argv[3], to, argv[5],
            [](std::string n, file_type_e t, std::vector<std::complex<int16_t>> &v) { return read_complex_data_from_file(n, t, v); },
            [](std::string n, file_type_e t, std::vector<std::complex<int16_t>> &v) { return write_complex_data_to_file(n, t, v); });
    if (kind == "cf32")
        return convert<std::vector<std::complex<float>>>(
            ti, argv[3], to, argv[5],
            [](std::string n, file_type_e t, std::vector<std::complex<float>> &v) { return read_complex_data_from_file(n, t, v); },
            [](std::string n, file_type_e t, std::vector<std::complex<float>> &v) { return write_complex_data_to_file(n, t, v); });
    if (kind == "i16")
        return convert<std::vector<int16_t>>(
            ti, argv[3], to, argv[5],
            [](std::string n, file_type_e t, std::vector<int16_t> &v) { return read_data_from_file(n, t, v); },
            [](std::string n, file_type_e t, std::vector<int16_t> &v) { return write_data_to_file(n, t, v); });
    if (kind == "f32")
        return convert<std::vector<float>>(
            ti, argv[3], to, argv[5],
            [](std::string n, file_type_e t, std::vector<float> &v) { return read_data_from_file(n, t, v); },
            [](std::string n, file_type_e t, std::vector<float> &v) { return write_data_to_file(n, t, v); });
    return 1;
}
