// Necklace-orbit evaluation of Schur cofactor minors (necklace.hip).
#pragma once
#include <stdint.h>

#include "device_context.h"

namespace phip {

// Largest Schur order with a compiled necklace kernel (minor order 46).
constexpr int kNecklaceMaxOrder = 47;

// True for a prime order 5 <= n <= kNecklaceMaxOrder and an odd p < 2^31.
bool necklace_supported(int n, uint64_t p);

// perm of the (0,0) minor of S_n (entries omega^(jk), 1 <= j, k < n) mod p,
// for omega^n == 1 mod p, by necklace orbits.  Returns 0 or -1.
int necklace_minor_modp(DeviceContext* ctx, int n, uint64_t omega, uint64_t p, int64_t* minor_out);

// The necklace walk of an odd prime order 3 <= n <= kNecklaceMaxOrder with
// no terms formed: the number of orbits under rotation and complement its
// representatives make up and the number of sign words (of length n - 1)
// they stand for.  Returns 0 or -1.
int necklace_orbit_stats(DeviceContext* ctx, int n, uint64_t* representatives, uint64_t* words);

}  // namespace phip
