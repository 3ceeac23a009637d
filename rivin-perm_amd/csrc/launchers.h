// Host-side launch helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace phip {

// Orders served by the register-resident 31-bit Glynn kernels.  Smaller
// orders (too few states for a 16-state unrolled chunk) and larger ones use
// the generic kernels.
constexpr int kGlynnMinOrder = 5;
constexpr int kGlynnMaxOrder = 46;

// Words per table column: residues one per word (group 0) or exact row sums
// packed two 16-bit halves per word (groups 2..4).
__host__ __device__ constexpr int glynn_stride(int n, int group) { return group == 0 ? n : (n + 1) / 2; }

}  // namespace phip
