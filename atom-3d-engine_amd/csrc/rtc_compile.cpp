// rtc_compile.cpp — the hipRTC compile of one generated scene source against the device headers
// embedded in the library: shared by the library (in-process compiles) and mm3e_rtc_worker
// (compiles in a child process, so that several scenes compile at once).
#include <hip/hiprtc.h>

#include <string>
#include <vector>

#include "codegen.h"
#include "rtc.h"

namespace mm3e {

bool rtc_compile_source(const std::string& src, std::vector<char>& code, std::string& err) {
  std::vector<const char*> hs, hn;
  for (int i = 0; i < kEmbeddedHeaderCount; i++) {
    hs.push_back(kEmbeddedHeaders[i].text);
    hn.push_back(kEmbeddedHeaders[i].name);
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "mm3e_gen.hip", (int)hs.size(), hs.data(), hn.data()) !=
      HIPRTC_SUCCESS) {
    err = "hiprtcCreateProgram failed";
    return false;
  }
  const char* opts[] = {"--offload-arch=gfx950",
                        "-O3",
                        "-std=c++17",
                        "-ffp-contract=off",
                        "-fhip-fp32-correctly-rounded-divide-sqrt",
                        "-Xclang",
                        "-target-feature",
                        "-Xclang",
                        "-packed-fp32-ops"};
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    err = std::string("hipRTC compile of the generated scene kernel failed: ") + hiprtcGetErrorString(r) + "\n" +
          log.substr(0, 4000);
    hiprtcDestroyProgram(&prog);
    return false;
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  code.resize(n);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  return true;
}

}  // namespace mm3e
