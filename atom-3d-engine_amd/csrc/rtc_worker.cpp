// mm3e_rtc_worker — compiles one generated scene source into a code object, in a process of
// its own.  hipRTC compiles one program at a time per process, so the library's code-generation
// threads each run this worker to compile several scenes at once (and compile in process when
// it cannot be run).  Usage: mm3e_rtc_worker <source file> <code object file>; exit status 0 on
// success, 2 with the compiler log on stderr on failure.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "rtc.h"

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: mm3e_rtc_worker <source file> <code object file>\n");
    return 2;
  }
  std::ifstream in(argv[1], std::ios::binary);
  if (!in) {
    std::fprintf(stderr, "mm3e_rtc_worker: cannot read %s\n", argv[1]);
    return 2;
  }
  std::stringstream ss;
  ss << in.rdbuf();
  std::vector<char> code;
  std::string err;
  if (!mm3e::rtc_compile_source(ss.str(), code, err)) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 2;
  }
  std::ofstream out(argv[2], std::ios::binary | std::ios::trunc);
  out.write(code.data(), (std::streamsize)code.size());
  out.close();
  if (!out) {
    std::fprintf(stderr, "mm3e_rtc_worker: cannot write %s\n", argv[2]);
    return 2;
  }
  return 0;
}
