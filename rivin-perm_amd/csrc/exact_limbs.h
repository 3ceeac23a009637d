// Exact products of up to eight 31-bit magnitudes in radix-2^31 limbs, for
// the exact-accumulation Glynn sink (glynn31.h, ExactSink).
//
// A value is an array of limbs l[0..k), value = sum l[i] 2^(31 i), every
// limb in [0, 2^31).  Multiplying two such arrays is schoolbook with 64-bit
// column sums: each product of two limbs is below 2^62, so a column of up
// to four products plus the incoming carry (below 2^34) stays below 2^64 and
// one unsigned 64-bit word holds it; the column's limb is its low 31 bits
// and its carry the rest.  On gfx950 every multiply-add below is one
// v_mad_u64_u32 (limb_mad).
//
// Inputs below 2^31 keep every result in range with no further condition:
// a pair product is below 2^62 (two limbs), a product of four below 2^124
// (four limbs, the top one below 2^31) and of eight below 2^248 (eight
// limbs).  Host and device share this code; a CPU test checks it against
// Python integers.
#pragma once
#include <stdint.h>

#ifndef __HIPCC__
#define PHIP_LIMB_FN static inline
#else
#define PHIP_LIMB_FN __host__ __device__ __forceinline__
#endif

namespace phip {

constexpr uint32_t kLimbMask = 0x7fffffffu;

// a * b + c as one v_mad_u64_u32 on the device.  Written as asm: compiled
// from C the column chains below kept a 64-bit multiply by the known-zero
// high word of every carry and many register moves (about 190 VALU
// instructions per Glynn state against about 140 this way).
PHIP_LIMB_FN uint64_t limb_mad(uint32_t a, uint32_t b, uint64_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t r;
    uint64_t carry;
    asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(carry) : "v"(a), "v"(b), "v"(c));
    return r;
#else
    return (uint64_t)a * b + c;
#endif
}

PHIP_LIMB_FN uint64_t limb_mul(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t r;
    uint64_t carry;
    asm("v_mad_u64_u32 %0, %1, %2, %3, 0" : "=v"(r), "=s"(carry) : "v"(a), "v"(b));
    return r;
#else
    return (uint64_t)a * b;
#endif
}

// a * b for a, b < 2^31: two limbs.
PHIP_LIMB_FN void limb_mul_1x1(uint32_t a, uint32_t b, uint32_t (&out)[2]) {
    const uint64_t t = limb_mul(a, b);
    out[0] = (uint32_t)t & kLimbMask;
    out[1] = (uint32_t)(t >> 31);
}

// Two-limb values (below 2^62) to four limbs.
PHIP_LIMB_FN void limb_mul_2x2(const uint32_t (&a)[2], const uint32_t (&b)[2], uint32_t (&out)[4]) {
    uint64_t t = limb_mul(a[0], b[0]);
    out[0] = (uint32_t)t & kLimbMask;
    t = limb_mad(a[0], b[1], t >> 31);
    t = limb_mad(a[1], b[0], t);
    out[1] = (uint32_t)t & kLimbMask;
    t = limb_mad(a[1], b[1], t >> 31);
    out[2] = (uint32_t)t & kLimbMask;
    out[3] = (uint32_t)(t >> 31);
}

// Four-limb values (below 2^124) to eight limbs.
PHIP_LIMB_FN void limb_mul_4x4(const uint32_t (&a)[4], const uint32_t (&b)[4], uint32_t (&out)[8]) {
    uint64_t t = limb_mul(a[0], b[0]);
    out[0] = (uint32_t)t & kLimbMask;
    t = limb_mad(a[0], b[1], t >> 31);
    t = limb_mad(a[1], b[0], t);
    out[1] = (uint32_t)t & kLimbMask;
    t = limb_mad(a[0], b[2], t >> 31);
    t = limb_mad(a[1], b[1], t);
    t = limb_mad(a[2], b[0], t);
    out[2] = (uint32_t)t & kLimbMask;
    t = limb_mad(a[0], b[3], t >> 31);
    t = limb_mad(a[1], b[2], t);
    t = limb_mad(a[2], b[1], t);
    t = limb_mad(a[3], b[0], t);
    out[3] = (uint32_t)t & kLimbMask;
    t = limb_mad(a[1], b[3], t >> 31);
    t = limb_mad(a[2], b[2], t);
    t = limb_mad(a[3], b[1], t);
    out[4] = (uint32_t)t & kLimbMask;
    t = limb_mad(a[2], b[3], t >> 31);
    t = limb_mad(a[3], b[2], t);
    out[5] = (uint32_t)t & kLimbMask;
    t = limb_mad(a[3], b[3], t >> 31);
    out[6] = (uint32_t)t & kLimbMask;
    out[7] = (uint32_t)(t >> 31);
}

// |prod of m[0..8)| in eight limbs, for magnitudes m[i] < 2^31.
PHIP_LIMB_FN void limb_product8(const uint32_t (&m)[8], uint32_t (&out)[8]) {
    uint32_t p0[2], p1[2], p2[2], p3[2], x[4], y[4];
    limb_mul_1x1(m[0], m[1], p0);
    limb_mul_1x1(m[2], m[3], p1);
    limb_mul_1x1(m[4], m[5], p2);
    limb_mul_1x1(m[6], m[7], p3);
    limb_mul_2x2(p0, p1, x);
    limb_mul_2x2(p2, p3, y);
    limb_mul_4x4(x, y, out);
}

}  // namespace phip
