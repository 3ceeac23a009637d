// Fixed-sign Glynn for odd moduli at or above 2^31 and the orders outside
// the persistent engine's range (wide.hip).
#pragma once
#include <stdint.h>

#include <vector>

#include "device_context.h"

namespace phip {

constexpr int kWideMaxOrder = 64;  // exclusive

struct WideJob {
    int index;   // position in the caller's modulus list (and matrix index)
    uint64_t q;  // odd, 3 <= q < 2^63
};

// odd_residue[job.index] = perm(A_index) mod job.q for every job, with A_i
// at A + i * stride (row-major n x n).  Returns 0 or -1.
int run_wide_glynn(DeviceContext* ctx, const int64_t* A, int64_t stride, int n, const std::vector<WideJob>& jobs,
                   std::vector<uint64_t>& odd_residue);

}  // namespace phip
