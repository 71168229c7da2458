// Compiled fused passes (jit.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "plan.hpp"

namespace qm {

struct State;

// Kernels of a plan's passes (cached by structure) and the coefficients they read.
struct JitProgram {
  std::vector<void *> fns;             // hipFunction_t per pass
  std::vector<double> coef;            // all passes' coefficients
  std::vector<int64_t> coef_offset;    // first coefficient of each pass
  std::vector<int> resident;           // workgroups per CU of each pass kernel
  // Batched compiled passes (FusedShape::state_bits): per pass, for each
  // coefficient slot, the op of the pass and the index in its c[] the value
  // is copied from (op, index pairs), so each state's coefficient block is
  // formed from its own plan's ops.
  std::vector<std::vector<int32_t>> prov;
};

bool jit_supported(const FusedPlan &plan);  // no run-time resolved ops
// The code variant the generator uses by default for a precision (QM_JIT_VARIANT).
int jit_default_variant(int precision);
JitProgram jit_prepare(const FusedPlan &plan, int precision, int device);
// The plan's kernels if every one is already compiled (no compilation).
bool jit_lookup(const FusedPlan &plan, int precision, int device, JitProgram *out);
// Compiles the plan's missing kernels on a background thread (at most a few
// plans at once; a plan already queued is not queued again).
void jit_prepare_async(const FusedPlan &plan, int precision, int device);
// Plans in the background (make() may search) and compiles the plan's missing
// kernels; a job for the same structure key is not queued twice.
void jit_plan_async(const std::string &structure, std::function<FusedPlan()> make, int precision, int device);
// Waits until no background compilation is running.
void jit_drain();
// Launches pass `pass` of a plan on state s: in place (src == dst, a pass of
// a plan without lane layout) or from src to dst (DevPass::out_place).
void jit_launch(State *s, const FusedShape &shape, const JitProgram &jp, size_t pass,
                const double *d_coef, hipStream_t st, const void *src, void *dst);
// Launches pass `pass` of a batched compiled plan over `items` consecutive
// states at `states` (2^shape.state_bits amplitudes each); d_coef points at
// the pass's coefficient blocks ([items][slots of the pass]).
void jit_launch_batch(int compute_units, int precision, const FusedShape &shape, const JitProgram &jp,
                      size_t pass, const double *d_coef, int64_t items, void *states, hipStream_t st);
void jit_stats(int64_t *compiled, int64_t *hits, double *seconds);
// Generates and compiles every pass kernel without loading it (no device needed).
int jit_compile_check(const FusedPlan &plan, int precision);

}  // namespace qm
