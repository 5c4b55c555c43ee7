// rtc.h — run-time compilation of generated scene kernels (hipRTC) and a process-wide module
// cache keyed by source text and device.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace mm3e {

// Compile `src` for gfx950 with the same floating-point contract as the library's own kernels
// (-O3, no contraction, correctly rounded division/sqrt, no packed FP32), load it on `device`
// (must be current) and return its `mm3e_gen_render`.  Identical sources compile once.
bool rtc_kernel(const std::string& src, int device, hipFunction_t* fn, std::string& err);

}  // namespace mm3e
