// Internal entry to the general IIR filter for other stages of the library.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime.h>

#include "../../include/fdsp.h"

namespace fdsp {

// sample layouts the IIR reads and writes
constexpr uint32_t kSampleCf32 = 0;  // complex float
constexpr uint32_t kSampleF32 = 1;   // real float (imaginary part zero)
constexpr uint32_t kSampleCi16 = 2;  // complex int16 (outputs truncated toward zero, saturated)
constexpr uint32_t kSampleI16 = 3;   // real int16

// filter n samples of layout tin at d_in into layout tout at d_out on stream
// s, advancing the handle's stream state; the caller serialises calls. The
// kernels' per-segment state records and per-block flags go into `work`
// (work_bytes of device memory) when it is given, else into the handle's own
// buffers, grown as needed.
int iir_apply_typed(fdsp_iir_t h, uint32_t tin, const void *d_in, size_t n, uint32_t tout, void *d_out,
                    hipStream_t s, void *work = nullptr, size_t work_bytes = 0);

}  // namespace fdsp
