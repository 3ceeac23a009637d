// Residue arithmetic for the odd moduli the signed 31-bit engine does not
// take (q >= 2^31).  Three policies share one interface so the wide Glynn
// kernel is written once:
//
//   Mont31L odd q < 2^31, unsigned Montgomery with R = 2^32 and one lazy
//           reduction of a sum of two products (polynomial kernels).
//   Mont32  odd q < 2^32, unsigned Montgomery with R = 2^32.  The 32-bit
//           primes of the reference's device planner (q = 1 mod n below
//           2^32, perm_cuda.py:289-302) land here.
//   M61     q = 2^61 - 1, the Mersenne prime: t = H 2^61 + L == H + L, no
//           multiplication beyond the 122-bit product.
//   PM61    q = 2^61 - c with odd 3 <= c < 2^28: pseudo-Mersenne folding
//           (two folds through c), no Montgomery factor.
//   Mont64  any other odd q < 2^63, unsigned Montgomery with R = 2^64.
//
// Every policy keeps values in [0, q); `add` and `mul` take and return
// values in that range.  `kRBits` is the Montgomery exponent: a product of
// k values through `mul` carries R^-(k-1) with R = 2^kRBits.
//
// PM61's 122-bit product is written as 32x32 -> 64 schoolbook steps, so the
// host build (the CPU tests compile this header with g++) runs exactly the
// device arithmetic; Mont64 takes the exact high product on either side.
#pragma once
#include <stdint.h>

#include "phip_arith.h"

#if defined(__HIPCC__)
#define PHIP_HDM __host__ __device__ __forceinline__
#else
#define PHIP_HDM inline
#endif

namespace phip {

// Every policy also has fma2(a, b, c, d) = a b + c d with one Montgomery
// factor, the step of a polynomial product; Mont31L reduces the exact sum
// once, the others reduce each product.

// Unsigned Montgomery, R = 2^32, odd p < 2^31: a b + c d < 2 p^2 < p 2^32,
// so the sum of two products takes a single REDC.
struct Mont31L {
    typedef uint32_t T;
    static constexpr int kRBits = 32;
    uint32_t p;
    uint32_t np;  // -p^-1 mod 2^32

    PHIP_HDM T add(T a, T b) const {
        const T s = a + b;  // < 2^32
        return s >= p ? s - p : s;
    }
    PHIP_HDM T neg(T a) const { return a ? p - a : 0; }
    PHIP_HDM T redc(uint64_t t) const {  // t < p 2^32 < 2^63: no carry
        const uint32_t m = (uint32_t)t * np;
        const uint64_t r = (t + (uint64_t)m * p) >> 32;
        return (T)(r >= p ? r - p : r);
    }
    PHIP_HDM T mul(T a, T b) const { return redc((uint64_t)a * b); }
    PHIP_HDM T fma2(T a, T b, T c, T d) const { return redc((uint64_t)a * b + (uint64_t)c * d); }
};

struct Mont32 {
    typedef uint32_t T;
    static constexpr int kRBits = 32;
    uint32_t p;
    uint32_t np;  // -p^-1 mod 2^32

    PHIP_HDM T add(T a, T b) const {
        const T s = a + b;
        return (s < a || s >= p) ? s - p : s;
    }
    PHIP_HDM T neg(T a) const { return a ? p - a : 0; }
    // REDC(a b): a b + m p is a multiple of 2^32 below 2^65, so the carry of
    // the 64-bit sum is kept as bit 32 of the quotient (< 2 p).
    PHIP_HDM T mul(T a, T b) const {
        const uint64_t t = (uint64_t)a * b;
        const uint32_t m = (uint32_t)t * np;
        const uint64_t s = t + (uint64_t)m * p;
        const uint64_t r = (s >> 32) | ((uint64_t)(s < t) << 32);
        return (T)(r >= p ? r - p : r);
    }
    PHIP_HDM T fma2(T a, T b, T c, T d) const { return add(mul(a, b), mul(c, d)); }
};

// The exact 122-bit product of a, b < 2^61 as (hi, lo) 64-bit words.
PHIP_HD void mul_61x61(uint64_t a, uint64_t b, uint64_t& hi, uint64_t& lo) {
    const uint64_t a0 = (uint32_t)a, a1 = a >> 32, b0 = (uint32_t)b, b1 = b >> 32;
    const uint64_t p00 = a0 * b0;
    const uint64_t mid = a0 * b1 + a1 * b0;  // < 2^62
    lo = p00 + (mid << 32);
    hi = a1 * b1 + (mid >> 32) + (lo < p00);
}

struct M61 {
    typedef uint64_t T;
    static constexpr int kRBits = 0;
    static constexpr uint64_t kMask = (1ull << 61) - 1;
    uint64_t p;  // 2^61 - 1

    PHIP_HDM T add(T a, T b) const {
        const T s = a + b;  // < 2^62
        return s >= p ? s - p : s;
    }
    PHIP_HDM T neg(T a) const { return a ? p - a : 0; }
    PHIP_HDM T mul(T a, T b) const {
        uint64_t hi, lo;
        mul_61x61(a, b, hi, lo);
        uint64_t r = (lo & kMask) + ((hi << 3) | (lo >> 61));  // < 2^62
        r = (r & kMask) + (r >> 61);                          // <= 2^61
        return r >= p ? r - p : r;
    }
    PHIP_HDM T fma2(T a, T b, T c, T d) const { return add(mul(a, b), mul(c, d)); }
};

struct PM61 {
    typedef uint64_t T;
    static constexpr int kRBits = 0;
    static constexpr uint64_t kMask = (1ull << 61) - 1;
    uint64_t p;  // 2^61 - c
    uint32_t c;  // odd, < 2^28
    uint32_t pad;

    PHIP_HDM T add(T a, T b) const {
        const T s = a + b;  // < 2^62
        return s >= p ? s - p : s;
    }
    PHIP_HDM T neg(T a) const { return a ? p - a : 0; }
    // a b = H 2^61 + L == H c + L (mod p); H c = H2 2^61 + L2 again, with
    // H2 < 2^28; one last fold of the at most two carried bits.
    PHIP_HDM T mul(T a, T b) const {
        uint64_t hi, lo;
        mul_61x61(a, b, hi, lo);  // product < 2^122
        const uint64_t L = lo & kMask;
        const uint64_t H = (hi << 3) | (lo >> 61);  // < 2^61
        const uint64_t h0c = (uint64_t)(uint32_t)H * c;
        const uint64_t h1c = (H >> 32) * c;
        const uint64_t lo2 = h0c + (h1c << 32);
        const uint64_t hi2 = (h1c >> 32) + (lo2 < h0c);
        const uint64_t L2 = lo2 & kMask;
        const uint64_t H2 = (hi2 << 3) | (lo2 >> 61);  // < 2^28
        uint64_t r = L + L2 + H2 * c;                   // < 2^63
        r = (r & kMask) + (r >> 61) * c;                // < 2^61 + 3 c < 2 p
        return r >= p ? r - p : r;
    }
    PHIP_HDM T fma2(T a, T b, T c, T d) const { return add(mul(a, b), mul(c, d)); }
};

struct Mont64 {
    typedef uint64_t T;
    static constexpr int kRBits = 64;
    uint64_t p;
    uint64_t np;  // -p^-1 mod 2^64

    PHIP_HDM T add(T a, T b) const { return add_mod_u64(a, b, p); }
    PHIP_HDM T neg(T a) const { return a ? p - a : 0; }
    PHIP_HDM T mul(T a, T b) const { return mont63_mul(a, b, p, np); }
    PHIP_HDM T fma2(T a, T b, T c, T d) const { return add(mul(a, b), mul(c, d)); }
};

// Arithmetic class of an odd modulus q >= 3 (q < 2^63).
enum WideClass { kWideMont32 = 0, kWidePM61 = 1, kWideMont64 = 2, kWideM61 = 3 };

PHIP_HD bool is_pm61(uint64_t q) {
    const uint64_t c = (1ull << 61) - q;
    return q < (1ull << 61) && (c & 1) && c < (1ull << 28);
}

PHIP_HD int wide_class(uint64_t q) {
    if (q < (1ull << 32)) return kWideMont32;
    if (q == (1ull << 61) - 1) return kWideM61;
    if (is_pm61(q)) return kWidePM61;
    return kWideMont64;
}

static inline Mont31L make_mont31l(uint64_t q) {
    return Mont31L{(uint32_t)q, 0u - inverse_mod_2_32((uint32_t)q)};
}
static inline Mont32 make_mont32(uint64_t q) { return Mont32{(uint32_t)q, 0u - inverse_mod_2_32((uint32_t)q)}; }
static inline M61 make_m61(uint64_t q) { return M61{q}; }
static inline PM61 make_pm61(uint64_t q) { return PM61{q, (uint32_t)((1ull << 61) - q), 0u}; }
static inline Mont64 make_mont64(uint64_t q) { return Mont64{q, 0 - inverse_mod_2_64(q)}; }

#if defined(__HIPCC__)
// Read-only tables read through the constant address space: a uniform
// address then compiles to scalar loads.  The empty asm hides the pointer's
// provenance so the compiler keeps them scalar.
template <class T>
using ConstPtr = const __attribute__((address_space(4))) T*;

template <class T>
__device__ __forceinline__ ConstPtr<T> scalar_ptr(const T* p) {
    asm volatile("" : "+s"(p));
    return (ConstPtr<T>)p;
}
#endif

}  // namespace phip
