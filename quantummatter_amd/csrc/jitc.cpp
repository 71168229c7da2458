// qm_jitc: compiles generated pass kernels in a child process.
//
// hipRTC runs one compilation at a time per process, so libqmhip.so spreads a
// plan's kernels over several of these helpers.  Usage:
//   qm_jitc DIR I0 I1 ...
// compiles DIR/kI.hip to DIR/kI.co for each index I, or writes the compiler's
// message to DIR/kI.err.  The exit status is 0 when every kernel compiled.
// The helper touches no GPU.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

#include "rtc.hpp"

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: qm_jitc DIR INDEX...\n");
    return 2;
  }
  const std::string dir = argv[1];
  int status = 0;
  for (int a = 2; a < argc; a++) {
    const std::string base = dir + "/k" + argv[a];
    std::ifstream in(base + ".hip", std::ios::binary);
    std::stringstream ss;
    ss << in.rdbuf();
    std::string err;
    const std::vector<char> code = in ? qm::rtc_compile(ss.str(), err) : std::vector<char>{};
    if (!in) err = "source missing";
    if (code.empty()) {
      std::ofstream(base + ".err") << err;
      status = 1;
      continue;
    }
    // Written under a temporary name and renamed, so a reader never sees a partial file.
    {
      std::ofstream out(base + ".tmp", std::ios::binary);
      out.write(code.data(), (std::streamsize)code.size());
      if (!out) {
        std::ofstream(base + ".err") << "write failed";
        status = 1;
        continue;
      }
    }
    if (std::rename((base + ".tmp").c_str(), (base + ".co").c_str()) != 0) {
      std::ofstream(base + ".err") << "rename failed";
      status = 1;
    }
  }
  return status;
}
