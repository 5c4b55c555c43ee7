// rtc.h — run-time compilation of generated scene kernels (hipRTC) and a process-wide module
// cache keyed by source text and device.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace mm3e {

// Optional kernels of a generated module, loaded when the source defines them.
enum GenOpt : int { GEN_WF_FIRST = 0, GEN_WF_CONT, GEN_WF_SHADE, GEN_FILM, GEN_GBUFFER, GEN_N_OPT };

// Compile `src` for gfx950 with the same floating-point contract as the library's own kernels
// (-O3, no contraction, correctly rounded division/sqrt, no packed FP32), load it on `device`
// (must be current) and return its `mm3e_gen_render`, when fn_o is given its ordered-tiles
// variant `mm3e_gen_render_o` (nullptr if the source has none) and, when fn_opt is given, the
// optional kernels of GenOpt (queued-frame stages mm3e_gen_wf_first / _cont / _shade, film
// mm3e_gen_film, G-buffer mm3e_gen_gbuffer; nullptr each if absent).  Identical sources compile
// once.
bool rtc_kernel(const std::string& src, int device, hipFunction_t* fn, hipFunction_t* fn_o, std::string& err,
                hipFunction_t* fn_opt = nullptr);
// Compile (or find in the cache) a generated module and return its kernel `name`.
bool rtc_function(const std::string& src, int device, const char* name, hipFunction_t* fn, std::string& err);

}  // namespace mm3e
