// Exact products of 31-bit magnitudes in radix-2^31 limbs, for the
// exact-accumulation Glynn sink (glynn31.h, ExactSink): eight values in
// eight limbs, and up to sixteen values modulo 2^(31 L); and eight signed
// values in eight signed limbs (slimb_product8).
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
#define PHIP_LIMB_MEMBER static inline
#else
#define PHIP_LIMB_FN __host__ __device__ __forceinline__
#define PHIP_LIMB_MEMBER __host__ __device__ __forceinline__ static
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

// Signed limbs: value = sum l[i] 2^(31 i) with l[0..k-1) in [0, 2^31) and
// the top limb signed (the floor of value / 2^(31 (k-1))), so a product of
// signed values needs no magnitudes and no sign.  Each signed multiply-add
// is one v_mad_i64_i32 on the device, into a signed 64-bit column word that
// must stay within (-2^63, 2^63): a product of two limbs (top limbs signed)
// is below 2^62 in magnitude, so two of them plus a carry below 2^33 fit.
PHIP_LIMB_FN int64_t slimb_mad(int32_t a, int32_t b, int64_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    int64_t r;
    uint64_t carry;
    asm("v_mad_i64_i32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(carry) : "v"(a), "v"(b), "v"(c));
    return r;
#else
    return (int64_t)a * b + c;
#endif
}

PHIP_LIMB_FN int64_t slimb_mul(int32_t a, int32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    int64_t r;
    uint64_t carry;
    asm("v_mad_i64_i32 %0, %1, %2, %3, 0" : "=v"(r), "=s"(carry) : "v"(a), "v"(b));
    return r;
#else
    return (int64_t)a * b;
#endif
}

// a * b for |a|, |b| < 2^31: two signed limbs.
PHIP_LIMB_FN void slimb_mul_1x1(int32_t a, int32_t b, int32_t (&out)[2]) {
    const int64_t t = slimb_mul(a, b);
    out[0] = (int32_t)((uint32_t)t & kLimbMask);
    out[1] = (int32_t)(t >> 31);
}

// Two-limb signed values (|value| < 2^62) to four limbs: every column holds
// at most two products and a carry below 2^32.  The top limb is the floor of
// the product / 2^93, in [-2^31, 2^31) since |product| < 2^124.
PHIP_LIMB_FN void slimb_mul_2x2(const int32_t (&a)[2], const int32_t (&b)[2], int32_t (&out)[4]) {
    int64_t t = slimb_mul(a[0], b[0]);
    out[0] = (int32_t)((uint32_t)t & kLimbMask);
    t = slimb_mad(a[0], b[1], t >> 31);
    t = slimb_mad(a[1], b[0], t);
    out[1] = (int32_t)((uint32_t)t & kLimbMask);
    t = slimb_mad(a[1], b[1], t >> 31);
    out[2] = (int32_t)((uint32_t)t & kLimbMask);
    out[3] = (int32_t)(t >> 31);
}

// Four-limb signed values to eight limbs, for |a|, |b| < 2^122 (top limbs
// below 2^29 in magnitude; the planner checks the bound of every four-group
// product).  Columns of unsigned limbs only (0..2, and the two unsigned
// products of column 3) run unsigned as limb_mul_4x4 does; column 3 is split
// before its two signed products (each below 2^60), so every signed column
// word stays below 2^62 + 2^61 + 2^33 in magnitude.
PHIP_LIMB_FN void slimb_mul_4x4(const int32_t (&a)[4], const int32_t (&b)[4], int32_t (&out)[8]) {
    const uint32_t a0 = (uint32_t)a[0], a1 = (uint32_t)a[1], a2 = (uint32_t)a[2];
    const uint32_t b0 = (uint32_t)b[0], b1 = (uint32_t)b[1], b2 = (uint32_t)b[2];
    uint64_t u = limb_mul(a0, b0);
    out[0] = (int32_t)((uint32_t)u & kLimbMask);
    u = limb_mad(a0, b1, u >> 31);
    u = limb_mad(a1, b0, u);
    out[1] = (int32_t)((uint32_t)u & kLimbMask);
    u = limb_mad(a0, b2, u >> 31);
    u = limb_mad(a1, b1, u);
    u = limb_mad(a2, b0, u);
    out[2] = (int32_t)((uint32_t)u & kLimbMask);
    u = limb_mad(a1, b2, u >> 31);
    u = limb_mad(a2, b1, u);
    const int64_t spill = (int64_t)(u >> 31);
    int64_t t = (int64_t)((uint32_t)u & kLimbMask);
    t = slimb_mad(a[0], b[3], t);
    t = slimb_mad(a[3], b[0], t);
    out[3] = (int32_t)((uint32_t)t & kLimbMask);
    t = slimb_mad(a[1], b[3], (t >> 31) + spill);
    t = slimb_mad(a[3], b[1], t);
    t = slimb_mad(a[2], b[2], t);
    out[4] = (int32_t)((uint32_t)t & kLimbMask);
    t = slimb_mad(a[2], b[3], t >> 31);
    t = slimb_mad(a[3], b[2], t);
    out[5] = (int32_t)((uint32_t)t & kLimbMask);
    t = slimb_mad(a[3], b[3], t >> 31);
    out[6] = (int32_t)((uint32_t)t & kLimbMask);
    out[7] = (int32_t)(t >> 31);
}

// prod g[0..8) in eight signed limbs, for |g[i]| < 2^31 whose four-value
// products g[0..4) and g[4..8) are below 2^122 in magnitude.
PHIP_LIMB_FN void slimb_product8(const int32_t (&g)[8], int32_t (&out)[8]) {
    int32_t p0[2], p1[2], p2[2], p3[2], x[4], y[4];
    slimb_mul_1x1(g[0], g[1], p0);
    slimb_mul_1x1(g[2], g[3], p1);
    slimb_mul_1x1(g[4], g[5], p2);
    slimb_mul_1x1(g[6], g[7], p3);
    slimb_mul_2x2(p0, p1, x);
    slimb_mul_2x2(p2, p3, y);
    slimb_mul_4x4(x, y, out);
}

// (a * b) mod 2^(31 LO) in LO limbs, for a of LA limbs and b of LB limbs,
// LO <= LA + LB.  Column k is the sum of the limb products a_i b_j with
// i + j = k plus the carry out of column k - 1, in one unsigned 64-bit
// word t.  A limb product is at most (2^31 - 1)^2 = 2^62 - 2^32 + 1, so
// four of them fit on top of a word below 2^33 and three on top of one
// below 2^62.  A column whose products do not all fit is split: the low 31
// bits of t stay in t (four more products fit) and t >> 31 goes to the
// column's carry-out.  At most eight products per column (the shorter
// operand has at most eight limbs) means at most two splits, so a carry-out
// is a sum of at most three words shifted right by 31, below 3 * 2^33; a
// column entered with such a carry takes three products before its first
// split, one entered with the carry of an unsplit column (below 2^33) four.
// Every branch below depends on loop indices only and folds away once the
// loops are unrolled.  The top limb of a full product (LO = LA + LB) is its
// carry, below 2^31 because the product is below 2^(31 LO).
template <int LA, int LB, int LO>
PHIP_LIMB_FN void limb_mul_trunc(const uint32_t (&a)[LA], const uint32_t (&b)[LB], uint32_t (&out)[LO]) {
    static_assert(LO <= LA + LB, "more output limbs than the product has");
    static_assert(LA <= 8 || LB <= 8, "at most eight limb products per column");
    uint64_t carry = 0;
    bool carry_small = true;
#pragma unroll
    for (int k = 0; k < LO; ++k) {
        if (k == LA + LB - 1) {
            // No products in the top column of a full product.
            out[k] = (uint32_t)carry;
            break;
        }
        uint64_t t = carry;
        uint64_t spill = 0;
        int room = carry_small ? 4 : 3;
        bool split = false;
        bool empty = k == 0;
#pragma unroll
        for (int i = 0; i < LA; ++i) {
            const int j = k - i;
            if (j < 0 || j >= LB) continue;
            if (room == 0) {
                spill += t >> 31;
                t &= kLimbMask;
                room = 4;
                split = true;
            }
            t = empty ? limb_mul(a[i], b[j]) : limb_mad(a[i], b[j], t);
            empty = false;
            --room;
        }
        out[k] = (uint32_t)t & kLimbMask;
        carry = (t >> 31) + spill;
        carry_small = !split;
    }
}

// Product of m[Lo..Hi) (each below 2^31) modulo 2^(31 L).  Split at the
// largest power of two below the count, so eight values multiply as
// limb_product8 does.
//
// V is the width the caller guarantees for the partial products: a node of
// c >= 2 values keeps limbs_for(c, V) = ceil(c V / 31) limbs (at most L), so
// its truncated product is its exact product only when that product is
// below 2^(31 limbs_for(c, V)).  With V = 31 that holds for any values below
// 2^31 (c values, c limbs).  A smaller V (values of about V bits, such as
// the products of two row sums of a large-entry table) drops the limb
// columns that must be zero; the host checks every internal node's bound
// product against its limbs before it picks such a kernel
// (compact_nodes_fit in the batch planner).
constexpr int limb_split(int count) {
    int s = 1;
    while (2 * s < count) s *= 2;
    return s;
}

constexpr int limbs_for(int count, int V) { return count <= 1 ? 1 : (count * V + 30) / 31; }

template <int Lo, int Hi, int L, int V = 31>
struct LimbProduct {
    static constexpr int kCount = Hi - Lo;
    static constexpr int kLimbs = limbs_for(kCount, V) < L ? limbs_for(kCount, V) : L;
    template <int M>
    PHIP_LIMB_MEMBER void run(const uint32_t (&m)[M], uint32_t (&out)[kLimbs]) {
        static_assert(Hi <= M && kCount >= 1, "range outside the values");
        static_assert(V >= 1 && V <= 31, "values are below 2^31");
        if constexpr (kCount == 1) {
            out[0] = m[Lo];
        } else {
            constexpr int kMid = Lo + limb_split(kCount);
            using A = LimbProduct<Lo, kMid, L, V>;
            using B = LimbProduct<kMid, Hi, L, V>;
            uint32_t x[A::kLimbs], y[B::kLimbs];
            A::run(m, x);
            B::run(m, y);
            limb_mul_trunc(x, y, out);
        }
    }
};

}  // namespace phip
