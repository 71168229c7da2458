// hipRTC compilation of one generated pass kernel to a gfx950 code object.
// Shared by the library (in-process compiles) and qm_jitc (the helper that
// compiles in child processes), so both produce identical code objects.
#pragma once

#include <hip/hiprtc.h>

#include <cstdlib>
#include <string>
#include <vector>

namespace qm {

// -ffp-contract=off keeps every product the generator writes unfused except
// the explicit fma calls; packed FP32 instructions are not generated.
inline std::vector<char> rtc_compile(const std::string &src, std::string &err) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "qm_pass.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
    err = "hiprtcCreateProgram failed";
    return {};
  }
  const char *level = std::getenv("QM_JIT_OPT");  // tuning knob: optimisation level (default -O3)
  const char *opts[] = {"--offload-arch=gfx950", level && *level ? level : "-O3", "-ffp-contract=off", "-Xclang",
                        "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  std::vector<char> code;
  if (rc != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    err = std::string(hiprtcGetErrorString(rc)) + ": " + log.substr(0, 4000);
  } else {
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    code.resize(n);
    hiprtcGetCode(prog, code.data());
  }
  hiprtcDestroyProgram(&prog);
  return code;
}

}  // namespace qm
