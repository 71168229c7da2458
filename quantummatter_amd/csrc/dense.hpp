// Dense-block passes on the matrix cores (dense.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "plan.hpp"

namespace qm {

struct State;

// Applies one dense block to a state; d_matrix holds the block's real
// (2^(k+1))^2 matrix in device memory, as doubles for a complex128 state and
// as floats for a complex64 state.
void launch_dense(State *st, const DenseBlock &b, const void *d_matrix, hipStream_t stream);

}  // namespace qm
