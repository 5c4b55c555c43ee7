// rtc.cpp — compile (in mm3e_rtc_worker child processes, or in process) and module cache for
// generated scene kernels.
#include "rtc.h"

#include <dlfcn.h>
#include <spawn.h>
#include <sys/wait.h>
#include <unistd.h>

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

extern char** environ;

#include "codegen.h"

namespace mm3e {

namespace {

struct Entry {
  std::mutex mu;       // held while compiling, so concurrent requests for one source wait
  bool done = false;
  bool ok = false;
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  hipFunction_t fn_o = nullptr;  // the same kernel (both dispatch orders)
  hipFunction_t fn_opt[GEN_N_OPT] = {};
  std::string err;
};

std::mutex g_mu;
std::map<std::pair<int, std::string>, std::shared_ptr<Entry>> g_cache;

// The worker binary beside this library (mm3e_rtc_worker), or "" when it is missing or the
// environment sets MM3E_RTC_PROCESS=0.
std::string worker_path() {
  static const std::string path = []() -> std::string {
    if (const char* v = std::getenv("MM3E_RTC_PROCESS"))
      if (std::string(v) == "0" || std::string(v) == "off") return "";
    Dl_info info;
    if (!dladdr((void*)&worker_path, &info) || !info.dli_fname) return "";
    std::string dir = info.dli_fname;
    const size_t slash = dir.rfind('/');
    dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
    const std::string p = dir + "/mm3e_rtc_worker";
    return access(p.c_str(), X_OK) == 0 ? p : "";
  }();
  return path;
}

// Compile in a child process: source and code object pass through two temporary files.
// Returns false (err says why) when the worker is missing, cannot start or fails.
bool compile_in_worker(const std::string& src, std::vector<char>& code, std::string& err) {
  const std::string worker = worker_path();
  if (worker.empty()) {
    err = "no mm3e_rtc_worker";
    return false;
  }
  const char* td = std::getenv("TMPDIR");
  std::string tmpl = std::string(td && *td ? td : "/tmp") + "/mm3e_rtc_XXXXXX";
  std::vector<char> name(tmpl.begin(), tmpl.end());
  name.push_back('\0');
  const int fd = mkstemp(name.data());
  if (fd < 0) {
    err = "mkstemp failed";
    return false;
  }
  const std::string in_path = name.data(), out_path = in_path + ".co";
  bool ok = write(fd, src.data(), src.size()) == (ssize_t)src.size();
  close(fd);
  if (ok) {
    char* argv[] = {const_cast<char*>(worker.c_str()), const_cast<char*>(in_path.c_str()),
                    const_cast<char*>(out_path.c_str()), nullptr};
    pid_t pid = 0;
    ok = posix_spawn(&pid, worker.c_str(), nullptr, nullptr, argv, environ) == 0;
    int status = 0;
    while (ok && waitpid(pid, &status, 0) < 0) {
      if (errno != EINTR) ok = false;
    }
    ok = ok && WIFEXITED(status) && WEXITSTATUS(status) == 0;
    if (!ok) err = "mm3e_rtc_worker failed";
  } else {
    err = "cannot write the generated source";
  }
  if (ok) {
    std::ifstream f(out_path, std::ios::binary);
    code.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    ok = !code.empty();
    if (!ok) err = "mm3e_rtc_worker wrote no code object";
  }
  unlink(in_path.c_str());
  unlink(out_path.c_str());
  return ok;
}

// A compile that failed in the worker is repeated in process, which also gives the compiler
// log of a real compile error.
bool compile(const std::string& src, std::vector<char>& code, std::string& err) {
  std::string werr;
  if (compile_in_worker(src, code, werr)) return true;
  return rtc_compile_source(src, code, err);
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
  if (e.ok && src.find("mm3e_gen_render_o(") != std::string::npos) {
    // one kernel serves both dispatch orders (a null TileOrder::order is index order)
    if (hipModuleGetFunction(&e.fn_o, e.mod, "mm3e_gen_render_o") != hipSuccess) e.fn_o = nullptr;
    e.fn = e.fn_o;
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
    e->err = "generated module has no mm3e_gen_render_o";
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
