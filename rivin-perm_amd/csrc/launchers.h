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

// Prime sets whose primes all lie below this bound may pair consecutive
// Gray states (PrimeSetSink with kPaired): every operand of the tree's last
// product is a Montgomery output, |v| < 2^62 / 2^32 + p / 2 < 1.625 2^30,
// so the sum of two such products plus the reduction's m p stays below
// (2 * 2.65 + 2.5) 2^60 < 2^63.
constexpr uint32_t kPairedPrimeLimit = 1342177280u;  // 1.25 * 2^30

// Words per table column: residues one per word (group 0) or exact row sums
// packed two 16-bit halves per word (groups 2..4).
__host__ __device__ constexpr int glynn_stride(int n, int group) { return group == 0 ? n : (n + 1) / 2; }
// Words per column of the zero-skipping walk's tables: biased byte row sums
// of a 0/1 matrix, four per word.
__host__ __device__ constexpr int byte_stride(int n) { return (n + 3) / 4; }

}  // namespace phip
