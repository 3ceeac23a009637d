// Coefficient-valued Glynn for per(C + x L) modulo odd moduli (affine.hip).
#pragma once
#include <stdint.h>

#include "device_context.h"

namespace phip {

// Largest order for moduli below 2^30 (the signed lazy arithmetic, tiers
// 48, 56 and 64 above order 40), for moduli below 2^32, and for moduli of
// 2^32 and above (64-bit residues: twice the registers per row and
// coefficient).
constexpr int kAffineMaxOrderS30 = 63;
constexpr int kAffineMaxOrder = 40;
constexpr int kAffineMaxOrderWide = 32;

struct AffineJob {
    int index;   // position in the caller's modulus list (and matrix pair index)
    uint64_t q;  // odd, 3 <= q < 2^63
};

// out[i * (n + 1) + d] = coefficient of x^d in per(C_i + x L_i) mod
// moduli[i], with C_i = C + i * stride and L_i = L + i * stride (row-major
// n x n int64).  0 <= n <= 63 for moduli below 2^30, n <= 40 below 2^32,
// n <= 32 above; odd moduli 3 <= m < 2^63.  min_tier forces a larger
// register tier (8..40 a multiple of 4, or 48, 56, 64 for moduli below
// 2^30; 64-bit residues round it up to a multiple of 8, at most 32).
// Returns 0 or -1.
int affine_permanent_residues(DeviceContext* ctx, const int64_t* C, const int64_t* L, int64_t stride, int n,
                              const int64_t* moduli, int k, int min_tier, int64_t* out);

}  // namespace phip
