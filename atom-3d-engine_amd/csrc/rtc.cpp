// rtc.cpp — hipRTC compile + module cache for generated scene kernels.
#include "rtc.h"

#include <hip/hiprtc.h>

#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "codegen.h"

namespace mm3e {

namespace {

struct Entry {
  std::mutex mu;       // held while compiling, so concurrent requests for one source wait
  bool done = false;
  bool ok = false;
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  hipFunction_t fn_o = nullptr;  // the ordered-tiles variant
  hipFunction_t fn_opt[GEN_N_OPT] = {};
  std::string err;
};

std::mutex g_mu;
std::map<std::pair<int, std::string>, std::shared_ptr<Entry>> g_cache;

bool compile(const std::string& src, std::vector<char>& code, std::string& err) {
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

}  // namespace

namespace {

std::shared_ptr<Entry> entry_for(const std::string& src, int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto& slot = g_cache[{device, src}];
  if (!slot) slot = std::make_shared<Entry>();
  return slot;
}

// Compile and load (e.mu held); a scene module's kernels are looked up once here.
void build(Entry& e, const std::string& src) {
  std::vector<char> code;
  e.ok = compile(src, code, e.err);
  if (e.ok && hipModuleLoadData(&e.mod, code.data()) != hipSuccess) {
    e.ok = false;
    e.err = "hipModuleLoadData of the generated scene kernel failed";
  }
  if (e.ok && src.find("mm3e_gen_render(") != std::string::npos) {
    if (hipModuleGetFunction(&e.fn, e.mod, "mm3e_gen_render") != hipSuccess) e.fn = nullptr;
    if (hipModuleGetFunction(&e.fn_o, e.mod, "mm3e_gen_render_o") != hipSuccess) e.fn_o = nullptr;
    static const char* opt_names[GEN_N_OPT] = {"mm3e_gen_wf_first", "mm3e_gen_wf_cont", "mm3e_gen_wf_shade",
                                               "mm3e_gen_film", "mm3e_gen_gbuffer"};
    for (int k = 0; k < GEN_N_OPT; k++)
      if (src.find(opt_names[k]) == std::string::npos ||
          hipModuleGetFunction(&e.fn_opt[k], e.mod, opt_names[k]) != hipSuccess)
        e.fn_opt[k] = nullptr;
  }
  e.done = true;
}

}  // namespace

bool rtc_kernel(const std::string& src, int device, hipFunction_t* fn, hipFunction_t* fn_o, std::string& err,
                hipFunction_t* fn_opt) {
  std::shared_ptr<Entry> e = entry_for(src, device);
  std::lock_guard<std::mutex> lk(e->mu);
  if (!e->done) build(*e, src);
  if (e->ok && !e->fn) {
    e->ok = false;
    e->err = "generated module has no mm3e_gen_render";
  }
  if (!e->ok) {
    err = e->err;
    return false;
  }
  *fn = e->fn;
  if (fn_o) *fn_o = e->fn_o;
  if (fn_opt)
    for (int k = 0; k < GEN_N_OPT; k++) fn_opt[k] = e->fn_opt[k];
  return true;
}

bool rtc_function(const std::string& src, int device, const char* name, hipFunction_t* fn, std::string& err) {
  std::shared_ptr<Entry> e = entry_for(src, device);
  std::lock_guard<std::mutex> lk(e->mu);
  if (!e->done) build(*e, src);
  if (!e->ok) {
    err = e->err;
    return false;
  }
  if (hipModuleGetFunction(fn, e->mod, name) != hipSuccess) {
    err = std::string("generated module has no ") + name;
    return false;
  }
  return true;
}

}  // namespace mm3e
