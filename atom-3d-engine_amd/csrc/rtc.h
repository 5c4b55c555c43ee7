// rtc.h — run-time compilation of generated scene kernels (hipRTC) and a process-wide module
// cache keyed by source text and device.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace mm3e {

// The hipRTC compile of one generated source (the options below) into a code object.
bool rtc_compile_source(const std::string& src, std::vector<char>& code, std::string& err);

// Optional kernels of a generated module, loaded when the source defines them.
enum GenOpt : int { GEN_WF_FIRST = 0, GEN_WF_CONT, GEN_WF_SHADE, GEN_FILM, GEN_GBUFFER, GEN_N_OPT };

// Compile `src` for gfx950 with the same floating-point contract as the library's own kernels
// (-O3, no contraction, correctly rounded division/sqrt, no packed FP32) — in a child process
// running mm3e_rtc_worker (installed beside the library) when it can be started, else in this
// process — load it on `device` (must be current) and return its render kernel
// `mm3e_gen_render_o` in fn and, when given, fn_o and, when fn_opt is given, the
// optional kernels of GenOpt (queued-frame stages mm3e_gen_wf_first / _cont / _shade, film
// mm3e_gen_film, G-buffer mm3e_gen_gbuffer; nullptr each if absent).  Identical sources compile
// once.
bool rtc_kernel(const std::string& src, int device, hipFunction_t* fn, hipFunction_t* fn_o, std::string& err,
                hipFunction_t* fn_opt = nullptr);
// Compile (or find in the cache) a generated module and return its kernel `name`.
bool rtc_function(const std::string& src, int device, const char* name, hipFunction_t* fn, std::string& err);

}  // namespace mm3e
