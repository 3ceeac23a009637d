// Modular arithmetic for the permanent kernels (gfx950 device code and the
// host set-up that prepares its constants).
//
// Two residue arithmetics are used:
//
//  * Signed Montgomery, R = 2^32, odd p < 2^31.  Values live in int32 and
//    may be negative.  mont31_mul(a, b) returns r == a*b*R^-1 (mod p) with
//    |r| < |a*b|/2^32 + p/2, so a chain of products whose operands stay
//    below 2^31 in magnitude never leaves (-2^31, 2^31) and never needs a
//    conditional correction.  The multiply is two 32x32->64 products plus
//    one 32-bit low product (v_mad_i64_i32 / v_mul_lo_u32 on CDNA).
//
//  * Unsigned Montgomery, R = 2^64, odd p < 2^63, for the wide-modulus
//    entry points (the reference accepts any positive signed 64-bit modulus,
//    perm_c_omp.c:876-880).
//
// All arithmetic is exact integer arithmetic; nothing is done in floating
// point.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define PHIP_HD __host__ __device__ __forceinline__
#else
#define PHIP_HD static inline
#endif

namespace phip {

// ---------------------------------------------------------------------
// Signed Montgomery, R = 2^32.
// ---------------------------------------------------------------------

// -p^-1 is not needed in the signed form: m = lo(a*b) * p^-1 mod 2^32 gives
// a*b - m*p == 0 (mod 2^32), so the high word of the exact 64-bit
// difference is the reduced value and no carry has to be recovered.
PHIP_HD int32_t mont31_mul(int32_t a, int32_t b, int32_t p, uint32_t pinv) {
    int64_t t = (int64_t)a * (int64_t)b;
    int32_t m = (int32_t)((uint32_t)t * pinv);
    int64_t u = t - (int64_t)m * (int64_t)p;
    return (int32_t)(u >> 32);
}

// p^-1 mod 2^32 by Newton iteration (each step doubles the correct bits;
// p*p == 1 mod 8 gives three correct bits to start).
PHIP_HD uint32_t inverse_mod_2_32(uint32_t p) {
    uint32_t x = p;
    for (int i = 0; i < 5; ++i) x *= 2u - p * x;
    return x;
}

PHIP_HD uint64_t inverse_mod_2_64(uint64_t p) {
    uint64_t x = p;
    for (int i = 0; i < 6; ++i) x *= 2u - p * x;
    return x;
}

// ---------------------------------------------------------------------
// Unsigned Montgomery, R = 2^64, odd p < 2^63 (inputs and outputs in [0, p)).
// ---------------------------------------------------------------------

PHIP_HD uint64_t mulhi_u64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// REDC(a*b) with m = lo(a*b) * (-p^-1); p < 2^63 keeps a*b + m*p < 2^127
// so the carry from the low words is (lo != 0).
PHIP_HD uint64_t mont63_mul(uint64_t a, uint64_t b, uint64_t p, uint64_t nprime) {
    uint64_t lo = a * b;
    uint64_t hi = mulhi_u64(a, b);
    uint64_t m = lo * nprime;
    uint64_t mh = mulhi_u64(m, p);
    uint64_t r = hi + mh + (lo != 0 ? 1u : 0u);
    return r >= p ? r - p : r;
}

PHIP_HD uint64_t add_mod_u64(uint64_t a, uint64_t b, uint64_t p) {
    return a >= p - b ? a - (p - b) : a + b;
}

PHIP_HD uint64_t sub_mod_u64(uint64_t a, uint64_t b, uint64_t p) {
    return a >= b ? a - b : p - (b - a);
}

// ---------------------------------------------------------------------
// Host-side helpers (exact, 128-bit); only called from host code.
// ---------------------------------------------------------------------
static inline uint64_t host_mulmod(uint64_t a, uint64_t b, uint64_t m) {
    return (uint64_t)(((unsigned __int128)a * b) % m);
}

static inline uint64_t host_powmod(uint64_t base, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    base %= m;
    while (e) {
        if (e & 1u) r = host_mulmod(r, base, m);
        base = host_mulmod(base, base, m);
        e >>= 1;
    }
    return r;
}

static inline int64_t host_floor_mod(int64_t a, int64_t m) {
    int64_t r = a % m;
    return r < 0 ? r + m : r;
}

}  // namespace phip
