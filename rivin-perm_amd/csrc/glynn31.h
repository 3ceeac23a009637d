// Glynn Gray-code kernels for odd moduli p < 2^31 (gfx950).
//
// Algorithm (the reference's fixed-sign Glynn, perm_c_omp.c:510-639):
//     per(A) = 2^-(n-1) * sum_{d in {+-1}^(n-1) x {+1}} (prod_k d_k)
//                                 * prod_i (sum_j d_j a_ij)
// traversed in Gray-code order over the first n-1 column signs, so moving
// from state s-1 to s flips column j = ctz(s): the row sums change by
// -2*a_ij when the new Gray bit is 1 and by +2*a_ij when it is 0
// (perm_c_omp.c:525-533).
//
// Work decomposition.  One lane owns 2^c consecutive states starting at an
// aligned base (lane << c).  Inside the chunk the flipped column is ctz(t)
// for the local index t, identical for every lane of the wavefront, so the
// column deltas are read with scalar loads and the n running row sums live
// in VGPRs.  The only lane-dependent flip direction is the one at
// t = 2^(c-1), which reads bit c of the lane's base.
//
// Row-sum representations (template parameter G):
//   G = 0     row sums are residues in [0, p), one per VGPR; a flip adds or
//             subtracts d = 2a mod p and reduces with one unsigned min.
//   G = 2..6  "small" matrices whose centred residues have row bounds
//             S_i = sum_j |a_ij| < 2^15 and, after the host balances rows
//             into groups (row order does not change the permanent), every
//             group product bound below 2^31 and every partial product that
//             feeds a 24-bit multiply below 2^23:
//             row sums are the exact integers, two per VGPR as 16-bit
//             halves, so a flip is one v_pk_add_u16 / v_pk_sub_i16 per two
//             rows (wrapping 16-bit arithmetic is exact because every true
//             row sum lies in (-2^15, 2^15)).  Each group of G row sums is
//             multiplied exactly with v_mul_i32_i24 reading the halves
//             through SDWA word selects, before one Montgomery step, so a
//             state costs ceil(n/G)-1 Montgomery products instead of n-1.
// Products are combined with the signed Montgomery multiply in a balanced
// tree; the accumulated sum therefore carries R^-(groups-1), which the
// finalisation step removes together with 2^-(n-1).
//
// Prime sets.  Packed row sums are exact integers, so every prime whose
// centred table is the same (all primes above twice the largest entry) sees
// the same row sums and group products.  glynn_chunk_set walks the states
// once for K such primes and runs only the Montgomery tree per prime.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "batch.h"
#include "exact_limbs.h"
#include "phip_arith.h"
#include "launchers.h"
#include "phip_reduce.h"

namespace phip {


constexpr int kGlynnUnroll = 16;  // states per unrolled step; chunk bits >= 4

// Exact 24-bit signed multiply (v_mul_i32_i24, full rate).  Operands must
// lie in [-2^23, 2^23); the callers guarantee it from the row-sum bound.
__device__ __forceinline__ int32_t mul_i24(int32_t a, int32_t b) {
    int32_t r;
    asm("v_mul_i32_i24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// Signed Montgomery product with the addend form of the second multiply:
// u = a*b + m*(-p) is one v_mad_i64_i32, and its high word is the result.
__device__ __forceinline__ int32_t mont31(int32_t a, int32_t b, int32_t negp, uint32_t pinv) {
    int64_t t = (int64_t)a * (int64_t)b;
    int32_t m = (int32_t)((uint32_t)t * pinv);
    int64_t u = t + (int64_t)m * (int64_t)negp;
    return (int32_t)(u >> 32);
}

// Column pointer for the flips: the constant address space lets the
// compiler read the wave-uniform deltas with scalar loads into SGPRs, and
// the empty asm hides the pointer from loop-invariant code motion so each
// flip issues its own loads instead of pinning every column in registers.
typedef const __attribute__((address_space(4))) int32_t* ConstCol;
__device__ __forceinline__ ConstCol opaque_ptr(const int32_t* p) {
    asm volatile("" : "+s"(p));
    return (ConstCol)p;
}

typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));

// v_mul_i32_i24 on sign-extended 16-bit halves (S = 0: low, 1: high) or on a
// full dword (S = 2), through SDWA source selects.
template <int S0, int S1>
__device__ __forceinline__ int32_t mul_sel(uint32_t a, uint32_t b);

#define PHIP_MUL_SEL(S0, S1, OPS, SELS)                                                             \
    template <>                                                                                     \
    __device__ __forceinline__ int32_t mul_sel<S0, S1>(uint32_t a, uint32_t b) {                    \
        int32_t r;                                                                                  \
        asm("v_mul_i32_i24_sdwa %0, " OPS " dst_sel:DWORD dst_unused:UNUSED_PAD " SELS              \
            : "=v"(r)                                                                               \
            : "v"(a), "v"(b));                                                                      \
        return r;                                                                                   \
    }
PHIP_MUL_SEL(0, 0, "sext(%1), sext(%2)", "src0_sel:WORD_0 src1_sel:WORD_0")
PHIP_MUL_SEL(0, 1, "sext(%1), sext(%2)", "src0_sel:WORD_0 src1_sel:WORD_1")
PHIP_MUL_SEL(1, 0, "sext(%1), sext(%2)", "src0_sel:WORD_1 src1_sel:WORD_0")
PHIP_MUL_SEL(1, 1, "sext(%1), sext(%2)", "src0_sel:WORD_1 src1_sel:WORD_1")
PHIP_MUL_SEL(2, 0, "%1, sext(%2)", "src0_sel:DWORD src1_sel:WORD_0")
PHIP_MUL_SEL(2, 1, "%1, sext(%2)", "src0_sel:DWORD src1_sel:WORD_1")
#undef PHIP_MUL_SEL

__device__ __forceinline__ uint32_t as_u32(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ u16x2 as_u16x2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }

// Row sums as residues in [0, p), one per register (G = 0).
template <int NR>
struct WideRows {
    static constexpr int kStride = NR;  // table words per column
    static constexpr int kGroups = NR;
    using Word = int32_t;
    int32_t r[NR];

    __device__ __forceinline__ static Word word_of(int32_t v) { return v; }

    __device__ __forceinline__ static int32_t add(int32_t x, int32_t d, int32_t p) {
        uint32_t a = (uint32_t)x + (uint32_t)d;
        uint32_t b = a - (uint32_t)p;
        return (int32_t)(a < b ? a : b);
    }
    __device__ __forceinline__ static int32_t sub(int32_t x, int32_t d, int32_t p) {
        uint32_t a = (uint32_t)x - (uint32_t)d;
        uint32_t b = a + (uint32_t)p;
        return (int32_t)(a < b ? a : b);
    }
    __device__ __forceinline__ void load(const int32_t* __restrict__ t) {
#pragma unroll
        for (int i = 0; i < NR; ++i) r[i] = t[i];
    }
    __device__ __forceinline__ void sub_masked(const int32_t* __restrict__ col, int32_t mask, int32_t p) {
#pragma unroll
        for (int i = 0; i < NR; ++i) r[i] = sub(r[i], col[i] & mask, p);
    }
    template <bool SUB>
    __device__ __forceinline__ void flip(const Word (&d)[kStride], int32_t p) {
#pragma unroll
        for (int i = 0; i < NR; ++i) r[i] = SUB ? sub(r[i], d[i], p) : add(r[i], d[i], p);
    }
    template <bool SUB>
    __device__ __forceinline__ void flip_mem(const int32_t* __restrict__ col, int32_t p) {
        const ConstCol c = opaque_ptr(col);
#pragma unroll
        for (int i = 0; i < NR; ++i) r[i] = SUB ? sub(r[i], c[i], p) : add(r[i], c[i], p);
    }
    __device__ __forceinline__ void flip_lane(const int32_t* __restrict__ col, bool sub_dir, int32_t p) {
        const ConstCol c = opaque_ptr(col);
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const int32_t d = c[i];
            r[i] = add(r[i], sub_dir ? p - d : d, p);  // p - 0 = p still reduces to r
        }
    }
    // Product of the row sums times R^-(NR-1); |result| < 2^31.
    __device__ __forceinline__ int32_t term(int32_t negp, uint32_t pinv) const {
        int32_t v[NR];
#pragma unroll
        for (int i = 0; i < NR; ++i) v[i] = r[i];
#pragma unroll
        for (int w = 1; w < NR; w *= 2) {
#pragma unroll
            for (int i = 0; i + w < NR; i += 2 * w) v[i] = mont31(v[i], v[i + w], negp, pinv);
        }
        return v[0];
    }
};

// Exact small row sums, two per register as 16-bit halves (G = 2..4).
template <int NR, int G>
struct PackedRows {
    static constexpr int kStride = (NR + 1) / 2;
    static constexpr int kGroups = (NR + G - 1) / G;
    using Word = u16x2;
    u16x2 r[kStride];

    __device__ __forceinline__ static Word word_of(int32_t v) { return as_u16x2((uint32_t)v); }

    __device__ __forceinline__ void load(const int32_t* __restrict__ t) {
#pragma unroll
        for (int k = 0; k < kStride; ++k) r[k] = as_u16x2((uint32_t)t[k]);
    }
    __device__ __forceinline__ void sub_masked(const int32_t* __restrict__ col, int32_t mask, int32_t) {
#pragma unroll
        for (int k = 0; k < kStride; ++k) r[k] = r[k] - as_u16x2((uint32_t)(col[k] & mask));
    }
    template <bool SUB>
    __device__ __forceinline__ void flip(const Word (&d)[kStride], int32_t) {
#pragma unroll
        for (int k = 0; k < kStride; ++k) r[k] = SUB ? r[k] - d[k] : r[k] + d[k];
    }
    template <bool SUB>
    __device__ __forceinline__ void flip_mem(const int32_t* __restrict__ col, int32_t) {
        const ConstCol c = opaque_ptr(col);
#pragma unroll
        for (int k = 0; k < kStride; ++k) {
            const u16x2 d = as_u16x2((uint32_t)c[k]);
            r[k] = SUB ? r[k] - d : r[k] + d;
        }
    }
    __device__ __forceinline__ void flip_lane(const int32_t* __restrict__ col, bool sub_dir, int32_t) {
        const ConstCol c = opaque_ptr(col);
#pragma unroll
        for (int k = 0; k < kStride; ++k) {
            const u16x2 d = as_u16x2((uint32_t)c[k]);
            r[k] = sub_dir ? r[k] - d : r[k] + d;
        }
    }

    template <int ROW>
    __device__ __forceinline__ uint32_t word() const {
        return as_u32(r[ROW / 2]);
    }

    template <int ROW>
    __device__ __forceinline__ int32_t row_value() const {
        const uint32_t w = word<ROW>();
        return (ROW & 1) ? ((int32_t)w >> 16) : (int32_t)(int16_t)(w & 0xffffu);
    }

    // Exact product of `C` consecutive row sums starting at FIRST, as
    //   1: r0   2: r0 r1   3: (r0 r1) r2   4: (r0 r1)(r2 r3)
    //   5: ((r0 r1) r2)(r3 r4)   6: ((r0 r1) r2)((r3 r4) r5)
    // with every 24-bit multiply operand below 2^23 in magnitude and the
    // result below 2^31 (the host checks these bounds when it picks G).
    template <int FIRST, int C>
    __device__ __forceinline__ int32_t run_product() const {
        if constexpr (C == 1) {
            return row_value<FIRST>();
        } else if constexpr (C == 2) {
            return mul_sel<FIRST & 1, (FIRST + 1) & 1>(word<FIRST>(), word<FIRST + 1>());
        } else if constexpr (C == 3) {
            return mul_sel<2, (FIRST + 2) & 1>((uint32_t)run_product<FIRST, 2>(), word<FIRST + 2>());
        } else {
            constexpr int A = (C + 1) / 2;
            return mul_i24(run_product<FIRST, A>(), run_product<FIRST + A, C - A>());
        }
    }

    template <int GI>
    __device__ __forceinline__ int32_t group_product() const {
        constexpr int first = GI * G;
        constexpr int count = (NR - first < G) ? NR - first : G;
        return run_product<first, count>();
    }

    template <int... GI>
    __device__ __forceinline__ void groups(int32_t (&v)[kGroups], std::integer_sequence<int, GI...>) const {
        ((v[GI] = group_product<GI>()), ...);
    }

    // Product of the row sums times R^-(kGroups-1); |result| < 2^31.
    __device__ __forceinline__ int32_t term(int32_t negp, uint32_t pinv) const {
        int32_t v[kGroups];
        groups(v, std::make_integer_sequence<int, kGroups>{});
#pragma unroll
        for (int w = 1; w < kGroups; w *= 2) {
#pragma unroll
            for (int i = 0; i + w < kGroups; i += 2 * w) v[i] = mont31(v[i], v[i + w], negp, pinv);
        }
        return v[0];
    }
};

template <int NR, int G>
using GlynnRows = typename std::conditional<G == 0, WideRows<NR>, PackedRows<NR, G>>::type;

// v_mul_i32_i24 on sign-extended bytes (S = 0..3) or on a full dword
// (S = 4), through SDWA source selects.
template <int S0, int S1>
__device__ __forceinline__ int32_t mul_bsel(uint32_t a, uint32_t b);

#define PHIP_MUL_BSEL(S0, S1, OPS, SELS)                                                            \
    template <>                                                                                     \
    __device__ __forceinline__ int32_t mul_bsel<S0, S1>(uint32_t a, uint32_t b) {                   \
        int32_t r;                                                                                  \
        asm("v_mul_i32_i24_sdwa %0, " OPS " dst_sel:DWORD dst_unused:UNUSED_PAD " SELS              \
            : "=v"(r)                                                                               \
            : "v"(a), "v"(b));                                                                      \
        return r;                                                                                   \
    }
#define PHIP_MUL_BSEL_ROW(S0)                                                                        \
    PHIP_MUL_BSEL(S0, 0, "sext(%1), sext(%2)", "src0_sel:BYTE_" #S0 " src1_sel:BYTE_0")              \
    PHIP_MUL_BSEL(S0, 1, "sext(%1), sext(%2)", "src0_sel:BYTE_" #S0 " src1_sel:BYTE_1")              \
    PHIP_MUL_BSEL(S0, 2, "sext(%1), sext(%2)", "src0_sel:BYTE_" #S0 " src1_sel:BYTE_2")              \
    PHIP_MUL_BSEL(S0, 3, "sext(%1), sext(%2)", "src0_sel:BYTE_" #S0 " src1_sel:BYTE_3")
PHIP_MUL_BSEL_ROW(0)
PHIP_MUL_BSEL_ROW(1)
PHIP_MUL_BSEL_ROW(2)
PHIP_MUL_BSEL_ROW(3)
PHIP_MUL_BSEL(4, 0, "%1, sext(%2)", "src0_sel:DWORD src1_sel:BYTE_0")
PHIP_MUL_BSEL(4, 1, "%1, sext(%2)", "src0_sel:DWORD src1_sel:BYTE_1")
PHIP_MUL_BSEL(4, 2, "%1, sext(%2)", "src0_sel:DWORD src1_sel:BYTE_2")
PHIP_MUL_BSEL(4, 3, "%1, sext(%2)", "src0_sel:DWORD src1_sel:BYTE_3")
#undef PHIP_MUL_BSEL_ROW
#undef PHIP_MUL_BSEL

// Row sums of tables with entries in {0, 1} (the zero-skipping walk).  A
// row sum lies in [-d, d] for the row's degree d <= 63, so it is kept as the
// byte s + 64, in [1, 127], four rows per 32-bit word (row 4k + b in byte b
// of word k).  The flip deltas 2 a_ij are bytes 0 or 2 and every true row
// sum stays in range, so a flip is one 32-bit add or subtract per four rows
// with no carry or borrow between bytes (a full-rate v_add_u32 /
// v_sub_u32, against one v_pk_add_u16 per two rows for PackedRows).  The
// padding bytes of an order that is not a multiple of four hold 65 (a row
// sum of 1) with zero deltas.  Tables: build_byte_tables.
template <int NR>
struct ByteRows {
    static constexpr int kStride = (NR + 3) / 4;
    using Word = uint32_t;
    uint32_t r[kStride];

    __device__ __forceinline__ static Word word_of(int32_t v) { return (uint32_t)v; }
    __device__ __forceinline__ void load(const int32_t* __restrict__ t) {
#pragma unroll
        for (int k = 0; k < kStride; ++k) r[k] = (uint32_t)t[k];
    }
    __device__ __forceinline__ void sub_masked(const int32_t* __restrict__ col, int32_t mask, int32_t) {
#pragma unroll
        for (int k = 0; k < kStride; ++k) r[k] -= (uint32_t)(col[k] & mask);
    }
    template <bool SUB>
    __device__ __forceinline__ void flip(const Word (&d)[kStride], int32_t) {
#pragma unroll
        for (int k = 0; k < kStride; ++k) r[k] = SUB ? r[k] - d[k] : r[k] + d[k];
    }
    template <bool SUB>
    __device__ __forceinline__ void flip_mem(const int32_t* __restrict__ col, int32_t) {
        const ConstCol c = opaque_ptr(col);
#pragma unroll
        for (int k = 0; k < kStride; ++k) {
            const uint32_t d = (uint32_t)c[k];
            r[k] = SUB ? r[k] - d : r[k] + d;
        }
    }
    // Subtracting d from a word is adding its 32-bit negation: the word's
    // value changes by -d in either form, and no byte leaves [1, 127].
    __device__ __forceinline__ void flip_lane(const int32_t* __restrict__ col, bool sub_dir, int32_t) {
        const ConstCol c = opaque_ptr(col);
#pragma unroll
        for (int k = 0; k < kStride; ++k) {
            const uint32_t d = (uint32_t)c[k];
            r[k] += sub_dir ? 0u - d : d;
        }
    }
    // Nonzero in some bit 7 exactly when a row sum of the word is zero:
    // every byte of v = w ^ 0x40404040 is (s + 64) ^ 64, in [0, 127] and
    // zero only for s = 0.  v - 0x01010101 leaves every byte of a word
    // without zero bytes in [0, 126] (no borrows), and turns the lowest zero
    // byte of any other word into 0xff (no borrow reaches it).  One
    // v_xad_u32 per word (x1010 = 0xfefefeff, the 32-bit -0x01010101, in a
    // register: a gfx9 VALU instruction reads at most one scalar operand);
    // the callers OR the results and test 0x80808080.
    __device__ __forceinline__ static uint32_t zero_bits(uint32_t w, uint32_t x1010) {
        uint32_t r;
        asm("v_xad_u32 %0, %1, %2, %3" : "=v"(r) : "v"(w), "s"(0x40404040u), "v"(x1010));
        return r;
    }
};

// a ^ b ^ c as v_bitop3_b32 (truth table 0x96), at the rate of one plain
// 32-bit logic instruction.
__device__ __forceinline__ uint32_t xor3_fast(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// a | b | c as v_bitop3_b32 (truth table 0xfe), which issues at the rate of
// the plain 32-bit logic instructions, about 1.7x that of v_or3_b32.
__device__ __forceinline__ uint32_t or3_fast(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0xfe" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// The signed row sums of a ByteRows word as two's-complement bytes: byte
// b + 64 stays below 256 (b <= 127), and flipping bit 7 of it gives
// b - 64 modulo 256.
__device__ __forceinline__ uint32_t byte_rows_signed(uint32_t w) { return (w + 0x40404040u) ^ 0x80808080u; }

// Exact products of signed byte row sums (words from byte_rows_signed),
// grouped as PackedRows groups them: `C` consecutive rows from FIRST as
//   1: r0   2: r0 r1   3: (r0 r1) r2   4: (r0 r1)(r2 r3)
//   5: ((r0 r1) r2)(r3 r4)   6: ((r0 r1) r2)((r3 r4) r5)
// under the same operand limits (plan_rows checks them on the host).
template <int NR, int G>
struct ByteGroups {
    static constexpr int kWords = (NR + 3) / 4;
    static constexpr int kGroups = (NR + G - 1) / G;
    const uint32_t (&w)[kWords];

    template <int ROW>
    __device__ __forceinline__ int32_t row_value() const {
        return (int32_t)(int8_t)(uint8_t)(w[ROW / 4] >> (8 * (ROW & 3)));
    }
    template <int FIRST, int C>
    __device__ __forceinline__ int32_t run_product() const {
        if constexpr (C == 1) {
            return row_value<FIRST>();
        } else if constexpr (C == 2) {
            return mul_bsel<FIRST & 3, (FIRST + 1) & 3>(w[FIRST / 4], w[(FIRST + 1) / 4]);
        } else if constexpr (C == 3) {
            return mul_bsel<4, (FIRST + 2) & 3>((uint32_t)run_product<FIRST, 2>(), w[(FIRST + 2) / 4]);
        } else {
            constexpr int A = (C + 1) / 2;
            return mul_i24(run_product<FIRST, A>(), run_product<FIRST + A, C - A>());
        }
    }
    template <int GI>
    __device__ __forceinline__ int32_t group_product() const {
        constexpr int first = GI * G;
        constexpr int count = (NR - first < G) ? NR - first : G;
        return run_product<first, count>();
    }
    template <int... GI>
    __device__ __forceinline__ void groups(int32_t (&v)[kGroups], std::integer_sequence<int, GI...>) const {
        ((v[GI] = group_product<GI>()), ...);
    }
};

// acc + S * v (S = +1 or -1) as one v_mad_i64_i32 with an inline constant.
// Written as asm so the accumulation stays in state order: left to itself
// the compiler reassociates the unrolled states' additions into a tree and
// keeps every term of every prime live at once, which exhausts the register
// file.
template <int S>
__device__ __forceinline__ int64_t mac_sign(int64_t acc, int32_t v) {
    int64_t r;
    uint64_t carry;
    if constexpr (S > 0) {
        asm("v_mad_i64_i32 %0, %1, %2, 1, %3" : "=v"(r), "=s"(carry) : "v"(v), "v"(acc));
    } else {
        asm("v_mad_i64_i32 %0, %1, %2, -1, %3" : "=v"(r), "=s"(carry) : "v"(v), "v"(acc));
    }
    return r;
}

// Per-state consumers of the Gray walk.  state<ODD>(rows) receives the row
// sums of every state of the chunk; ODD marks odd local indices, whose
// terms carry the opposite sign of the chunk's first state.
//
// One prime: the product of the row sums through the balanced Montgomery
// tree, added for even and subtracted for odd local indices.
template <int NR, int G>
struct OnePrimeSink {
    using Rows = GlynnRows<NR, G>;
    int32_t negp;
    uint32_t pinv;
    int64_t acc = 0;
    template <bool ODD>
    __device__ __forceinline__ void state(const Rows& rs) {
        acc = mac_sign<ODD ? -1 : 1>(acc, rs.term(negp, pinv));
    }
};

// Montgomery reduction of an exact signed product |t| < 2^62:
// t * R^-1 mod p with |result| < 2^31 (the second half of mont31).
__device__ __forceinline__ int32_t redc31(int64_t t, int32_t negp, uint32_t pinv) {
    int32_t m = (int32_t)((uint32_t)t * pinv);
    int64_t u = t + (int64_t)m * (int64_t)negp;
    return (int32_t)(u >> 32);
}

// acc + a * b as one in-order v_mad_i64_i32 (exact while the sum fits).
__device__ __forceinline__ int64_t mad_exact(int64_t acc, int32_t a, int32_t b) {
    int64_t r;
    uint64_t carry;
    asm("v_mad_i64_i32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(carry) : "v"(a), "v"(b), "v"(acc));
    return r;
}

// K primes sharing one packed table.  The exact row sums, the exact group
// products and the exact 64-bit products of adjacent groups (|g| < 2^31, so
// |g g'| < 2^62) do not depend on the prime and are formed once per state;
// each prime then reduces the pair products with one Montgomery reduction
// apiece (the first level of the same balanced tree, bit-identical to
// mont31 on the two groups) and combines the leaves with mont31.  The term
// therefore carries R^-(groups-1) exactly as in OnePrimeSink.
//
// kPaired (every prime below kPairedPrimeLimit): consecutive states 2i and
// 2i+1 carry opposite signs, so the last product of the tree is kept exact
// for the even state (with the first group negated, once for all primes)
// and the odd state's exact product is added to it; one Montgomery
// reduction of that difference replaces two products and two reductions.
// The term carries the same R^-(groups-1).  Shapes whose last product takes
// the unreduced odd group directly keep the per-state form.
template <int NR, int G, int K, bool kPaired = false>
struct PrimeSetSink {
    static_assert(G > 0, "prime sets need exact packed row sums");
    using Rows = PackedRows<NR, G>;
    static constexpr int kGroups = Rows::kGroups;
    static constexpr int kPairs = kGroups / 2;
    static constexpr int kLeaves = kPairs + (kGroups & 1);
    static constexpr int top_of(int leaves) {
        int w = 1;
        while (2 * w < leaves) w *= 2;
        return w;
    }
    // Index of the last product's right operand.
    static constexpr int kTop = top_of(kLeaves);
    static constexpr bool kPairs2 =
        kPaired && kLeaves >= 2 && !(kTop == kLeaves - 1 && (kGroups & 1));
    int32_t negp[K];
    uint32_t pinv[K];
    int64_t acc[K];
    int64_t held[kPairs2 ? K : 1];

    template <bool ODD>
    __device__ __forceinline__ void state(const Rows& rs) {
        int32_t g[kGroups];
        rs.groups(g, std::make_integer_sequence<int, kGroups>{});
        int64_t P[kPairs > 0 ? kPairs : 1];
        if constexpr (kPairs2 && !ODD) g[0] = -g[0];
#pragma unroll
        for (int j = 0; j < kPairs; ++j) P[j] = (int64_t)g[2 * j] * (int64_t)g[2 * j + 1];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            int32_t v[kLeaves];
#pragma unroll
            for (int j = 0; j < kPairs; ++j) v[j] = redc31(P[j], negp[k], pinv[k]);
            if constexpr (kGroups & 1) v[kLeaves - 1] = g[kGroups - 1];
            if constexpr (kPairs2) {
#pragma unroll
                for (int w = 1; w < kTop; w *= 2) {
#pragma unroll
                    for (int i = 0; i + w < kLeaves; i += 2 * w) v[i] = mont31(v[i], v[i + w], negp[k], pinv[k]);
                }
                if constexpr (!ODD) {
                    held[k] = mad_exact(0, v[0], v[kTop]);
                } else {
                    // held = -E(2i); -(E(2i) - E(2i+1)) reduced, subtracted.
                    const int64_t t = mad_exact(held[k], v[0], v[kTop]);
                    acc[k] = mac_sign<-1>(acc[k], redc31(t, negp[k], pinv[k]));
                }
            } else {
#pragma unroll
                for (int w = 1; w < kLeaves; w *= 2) {
#pragma unroll
                    for (int i = 0; i + w < kLeaves; i += 2 * w) v[i] = mont31(v[i], v[i + w], negp[k], pinv[k]);
                }
                acc[k] = mac_sign<ODD ? -1 : 1>(acc[k], v[0]);
            }
        }
    }
};

// Exact accumulation (any number of primes): each state's product of the
// exact group products (|g| < 2^31) is formed exactly, as a magnitude in
// radix-2^31 limbs (exact_limbs.h) and a sign, and every limb is added with
// that sign to its own int64 column accumulator.  No modular arithmetic
// runs in the walk: the lane's columns hold its exact signed sum of terms,
// sum_c acc[c] 2^(31 c), and any prime's residue follows from them at the
// end.
//
// L = 8 takes at most eight groups: the product of eight values below 2^31
// is below 2^248, eight limbs, with no further condition.  Per state that
// is the exact pair products (4 multiplies), the two 2 x 2 and one 4 x 4
// limb products (24 multiplies and 13 limb splits), the magnitudes and the
// sign, and 8 accumulations, for every prime at once.
//
// L = 10 or 12 takes up to sixteen groups, whose product can need sixteen
// limbs; the walk keeps it modulo 2^(31 L) (LimbProduct), which is the
// exact product whenever the product of the table's row bounds (sum_j
// |a_ij|, each bounding |row sum| in every state) is below 2^(31 L).  The
// planner sends a table here only then.
//
// V < 31 (the compact classes) sizes the product tree's internal nodes for
// values of about V bits (LimbProduct); the planner sends a table there only
// when every node's bound product fits its limbs.
//
// V = kSignedLimbs (L = 8) multiplies the signed group products in signed
// limbs (slimb_product8): no magnitudes and no sign per state, for tables
// whose four-group products are below 2^122 (the planner checks their bound
// products).
//
// Every limb is below 2^31 in magnitude, so after 2^cbits states
// |acc[c]| < 2^(31 + cbits).
template <int NR, int G, int L, int V = 31>
struct ExactSink {
    static_assert(G > 0, "exact accumulation needs exact packed row sums");
    static_assert(V != kSignedLimbs || L == 8, "signed limbs take eight groups in eight columns");
    using Rows = PackedRows<NR, G>;
    static constexpr int kGroups = Rows::kGroups;
    static_assert(L == 8 ? kGroups <= 8 : kGroups <= 16, "too many groups for the exact sink");
    int64_t acc[L];

    template <bool ODD>
    __device__ __forceinline__ void state(const Rows& rs) {
        int32_t g[kGroups];
        rs.groups(g, std::make_integer_sequence<int, kGroups>{});
        if constexpr (V == kSignedLimbs) {
            // Signed limbs: the group products as they are, the term's sign
            // only the state's.
            int32_t gv[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) gv[k] = k < kGroups ? g[k] : 1;
            int32_t z[8];
            slimb_product8(gv, z);
#pragma unroll
            for (int c = 0; c < 8; ++c) acc[c] = mac_sign<ODD ? -1 : 1>(acc[c], z[c]);
            return;
        }
        constexpr int kValues = L == 8 ? 8 : kGroups;
        uint32_t m[kValues];
#pragma unroll
        for (int k = 0; k < kValues; ++k) m[k] = k < kGroups ? (uint32_t)(g[k] < 0 ? -g[k] : g[k]) : 1u;
        // The term's sign: bit 31 of the XOR of the group products (and of
        // 0x80000000 for an odd state), folded three at a time.
        constexpr int kTerms = kGroups + (ODD ? 1 : 0);
        uint32_t t[kTerms];
#pragma unroll
        for (int k = 0; k < kGroups; ++k) t[k] = (uint32_t)g[k];
        if constexpr (ODD) t[kGroups] = 0x80000000u;
        uint32_t sign = t[0];
#pragma unroll
        for (int k = 1; k < kTerms; k += 2) sign = k + 1 < kTerms ? xor3_fast(sign, t[k], t[k + 1]) : sign ^ t[k];
        const int32_t e = ((int32_t)sign >> 31) | 1;
        if constexpr (L == 8) {
            uint32_t z[8];
            limb_product8(m, z);
#pragma unroll
            for (int c = 0; c < 8; ++c) acc[c] = mad_exact(acc[c], (int32_t)z[c], e);
        } else {
            using P = LimbProduct<0, kGroups, L, V>;
            uint32_t z[P::kLimbs];
            P::run(m, z);
#pragma unroll
            for (int c = 0; c < P::kLimbs; ++c) acc[c] = mad_exact(acc[c], (int32_t)z[c], e);
        }
    }
};

// K primes sharing one table of a 0/1 matrix, whose states mostly have a
// zero row sum (a row of even degree d sums to zero in C(d, d/2) / 2^d of
// the states, and one zero row zeroes the term).  The walk keeps the row
// sums as bytes (ByteRows) and tests each state with one v_xad_u32 per four
// rows and an AND over the words.  The group products and the per-prime
// trees run only for the other states: each wavefront queues them in LDS
// (their row words and the state's sign) and, once 64 are queued, every
// lane takes one, turns its words into signed bytes, forms its group
// products (ByteGroups, SDWA byte selects) and runs the trees of
// PrimeSetSink's per-state form at full lane occupancy.  A lane adds the
// terms of whichever states it takes, so the sign is absolute (the chunk's
// first-state parity is known before the walk) and the lane's sum is not
// negated at the end.  The queue holds kZeroQueue entries per row word plus
// one for the sign, column-major: between states it holds fewer than 64,
// and one state adds at most 64.
constexpr int kZeroQueue = 128;

//
// kHalfTest: the table's rows of even degree all lie in the leading
// (kWords + 1) / 2 words, and the test reads only those; a row of odd
// degree sums to an odd number in every state.
template <int NR, int G, int K, bool kHalfTest>
struct ZeroSkipSink {
    static_assert(G > 0, "prime sets need exact row sums");
    using Rows = ByteRows<NR>;
    using Groups = ByteGroups<NR, G>;
    static constexpr int kGroups = Groups::kGroups;
    static constexpr int kWords = Rows::kStride;
    static constexpr int kTestWords = kHalfTest ? (kWords + 1) / 2 : kWords;
    static constexpr int kPairs = kGroups / 2;
    static constexpr int kLeaves = kPairs + (kGroups & 1);
    int32_t negp[K];
    uint32_t pinv[K];
    int64_t acc[K];
    int32_t* queue;   // this wavefront's (kWords + 1) * kZeroQueue words of LDS
    uint32_t count;   // queued states (wave-uniform)
    uint32_t lane;
    bool neg;         // the chunk's first state has sign -1
    bool active;      // the lane's chunk belongs to the job
    uint64_t active_mask;  // the wavefront's active lanes
    uint32_t x1010;   // 0xfefefeff (ByteRows::zero_bits)

    // Adds the term of queued entry `at`: its group products from its row
    // words, the first one negated for a state of sign -1.
    __device__ __forceinline__ void consume(uint32_t at) {
        uint32_t w[kWords];
#pragma unroll
        for (int k = 0; k < kWords; ++k) w[k] = byte_rows_signed((uint32_t)queue[k * kZeroQueue + at]);
        int32_t g[kGroups];
        const Groups q{w};
        q.groups(g, std::make_integer_sequence<int, kGroups>{});
        if (queue[kWords * kZeroQueue + at]) g[0] = -g[0];
        int64_t P[kPairs > 0 ? kPairs : 1];
#pragma unroll
        for (int j = 0; j < kPairs; ++j) P[j] = (int64_t)g[2 * j] * (int64_t)g[2 * j + 1];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            int32_t v[kLeaves];
#pragma unroll
            for (int j = 0; j < kPairs; ++j) v[j] = redc31(P[j], negp[k], pinv[k]);
            if constexpr (kGroups & 1) v[kLeaves - 1] = g[kGroups - 1];
#pragma unroll
            for (int wd = 1; wd < kLeaves; wd *= 2) {
#pragma unroll
                for (int i = 0; i + wd < kLeaves; i += 2 * wd) v[i] = mont31(v[i], v[i + wd], negp[k], pinv[k]);
            }
            acc[k] = mac_sign<1>(acc[k], v[0]);
        }
    }

    template <bool ODD>
    __device__ __forceinline__ void state(const Rows& rs) {
        uint32_t m = Rows::zero_bits(rs.r[0], x1010);
#pragma unroll
        for (int k = 1; k < kTestWords; k += 2) {
            if (k + 1 < kTestWords) {
                m = or3_fast(m, Rows::zero_bits(rs.r[k], x1010), Rows::zero_bits(rs.r[k + 1], x1010));
            } else {
                m |= Rows::zero_bits(rs.r[k], x1010);
            }
        }
        // The ballot of the bare compare is its lane mask; with `active`
        // folded in first the compiler rebuilt the mask from a VGPR copy of
        // the boolean every state.
        const bool zero_free = (m & 0x80808080u) == 0u;
        const uint64_t mask = __ballot(zero_free) & active_mask;
        if (mask == 0) return;
        if (zero_free && active) {
            const uint32_t pos =
                count + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
#pragma unroll
            for (int k = 0; k < kWords; ++k) queue[k * kZeroQueue + pos] = (int32_t)rs.r[k];
            queue[kWords * kZeroQueue + pos] = (ODD != neg) ? 1 : 0;
        }
        count += (uint32_t)__popcll(mask);
        if (count >= 64) {
            count -= 64;
            // Other lanes' queue writes above are read below; LDS keeps a
            // wavefront's accesses in order, and this keeps the compiler
            // from moving the reads up.
            __builtin_amdgcn_wave_barrier();
            consume(count + lane);
        }
    }

    // Takes the states still queued (fewer than 64), one per lane.
    __device__ __forceinline__ void flush() {
        if (count == 0) return;
        __builtin_amdgcn_wave_barrier();
        if (lane < count) consume(lane);
        count = 0;
    }
};

// Walks one lane's chunk: the 2^cbits Gray states starting at `base` (a
// multiple of 2^cbits), handing the row sums of each state to the sink.
// `tables` holds the row sums of state 0 (all column signs +1, in the Rows
// representation: exact centred sums packed two per word, biased bytes four
// per word, or residues) followed by the flip deltas of the n-1 Gray
// columns, column-major with Rows::kStride words per column (2a_ij in the
// same representation, or 2a_ij mod p; p is read only by WideRows).
// Returns the parity of gray(base), the sign of the chunk's first state.
template <int NR, class Rows, class Sink>
__device__ __forceinline__ bool gray_walk_rows(const int32_t* __restrict__ tables, int n, int cbits, uint64_t base,
                                               int32_t p, Sink& sink) {
    constexpr int W = Rows::kStride;
    const int32_t* __restrict__ D = tables + W;
    const uint64_t gray = base ^ (base >> 1);

    // Row sums of the chunk's first state: subtract 2a for every set Gray
    // bit (perm_c_omp.c:582-600).  Bits below c-1 of gray(base) are zero.
    Rows rs;
    rs.load(tables);
    for (int j = cbits - 1; j < n - 1; ++j) {
        rs.sub_masked(D + j * W, -(int32_t)((gray >> j) & 1u), p);
    }

    // Column 0 flips every other state; keep it resident.
    typename Rows::Word c0[W];
#pragma unroll
    for (int k = 0; k < W; ++k) c0[k] = Rows::word_of(D[k]);
    const int32_t* __restrict__ d1 = D + W;
    const int32_t* __restrict__ d2 = D + 2 * W;

    sink.template state<false>(rs);
    const uint32_t T = 1u << cbits;
    for (uint32_t t0 = 0; t0 < T; t0 += kGlynnUnroll) {
        // Directions for s = 1..15 follow from bit j+1 of t0+s, which lies
        // inside the unrolled block except for s = 8 (bit 4 of base+t0).
        rs.template flip<true>(c0, p);         sink.template state<true>(rs);   // s=1
        rs.template flip_mem<true>(d1, p);     sink.template state<false>(rs);  // s=2
        rs.template flip<false>(c0, p);        sink.template state<true>(rs);   // s=3
        rs.template flip_mem<true>(d2, p);     sink.template state<false>(rs);  // s=4
        rs.template flip<true>(c0, p);         sink.template state<true>(rs);   // s=5
        rs.template flip_mem<false>(d1, p);    sink.template state<false>(rs);  // s=6
        rs.template flip<false>(c0, p);        sink.template state<true>(rs);   // s=7
        {
            // s=8 flips column 3; its new Gray bit is 1 ^ bit 4 of (base+t),
            // which is bit 4 of t, the same in every lane, when the chunk
            // has more than 4 bits.
            if (cbits > 4) {
                if (((t0 + 8) >> 4) & 1u) {
                    rs.template flip_mem<false>(D + 3 * W, p);
                } else {
                    rs.template flip_mem<true>(D + 3 * W, p);
                }
            } else {
                const uint64_t t = base + t0 + 8;
                rs.flip_lane(D + 3 * W, !((t >> 4) & 1u), p);
            }
            sink.template state<false>(rs);
        }
        rs.template flip<true>(c0, p);         sink.template state<true>(rs);   // s=9
        rs.template flip_mem<true>(d1, p);     sink.template state<false>(rs);  // s=10
        rs.template flip<false>(c0, p);        sink.template state<true>(rs);   // s=11
        rs.template flip_mem<false>(d2, p);    sink.template state<false>(rs);  // s=12
        rs.template flip<true>(c0, p);         sink.template state<true>(rs);   // s=13
        rs.template flip_mem<false>(d1, p);    sink.template state<false>(rs);  // s=14
        rs.template flip<false>(c0, p);        sink.template state<true>(rs);   // s=15
        const uint32_t tn = t0 + kGlynnUnroll;
        if (tn < T) {
            const int j = __builtin_ctz(tn);  // 4 <= j <= cbits-1, uniform
            // The new Gray bit is 1 ^ bit j+1 of (base+tn): bit j+1 of tn in
            // every lane except at the chunk's middle state (j = cbits-1),
            // where it is the lane's own base bit.
            if (j + 1 < cbits) {
                if ((tn >> (j + 1)) & 1u) {
                    rs.template flip_mem<false>(D + j * W, p);
                } else {
                    rs.template flip_mem<true>(D + j * W, p);
                }
            } else {
                const uint64_t t = base + tn;
                rs.flip_lane(D + j * W, !((t >> (j + 1)) & 1u), p);
            }
            sink.template state<false>(rs);
        }
    }
    return __popcll(gray) & 1;
}

// The walk over GlynnRows<NR, G> tables (glynn_stride(NR, G) words per
// column).
template <int NR, int G, class Sink>
__device__ __forceinline__ bool gray_walk(const int32_t* __restrict__ tables, int n, int cbits, uint64_t base,
                                          int32_t p, Sink& sink) {
    return gray_walk_rows<NR, GlynnRows<NR, G>>(tables, n, cbits, base, p, sink);
}

// The lane's signed sum of terms as one Montgomery residue: |acc| <
// 2^(31+cbits) <= 2^61, so redc31 applies and the result is congruent to
// acc * R^-1 with |result| < 2^31 (no 64-bit division in the epilogue).
__device__ __forceinline__ int32_t lane_value(int64_t acc, bool negate, int32_t negp, uint32_t pinv) {
    const int32_t r = redc31(acc, negp, pinv);
    return negate ? -r : r;
}

// One prime: the lane's value, congruent to its signed sum of terms times
// R^-groups; the finalisation multiplies by R^groups 2^-(n-1).
template <int NR, int G>
__device__ __forceinline__ int32_t glynn_chunk(const int32_t* __restrict__ tables, int n, int cbits, uint64_t base,
                                               int32_t p, uint32_t pinv) {
    OnePrimeSink<NR, G> sink;
    sink.negp = -p;
    sink.pinv = pinv;
    const bool neg = gray_walk<NR, G>(tables, n, cbits, base, p, sink);
    return lane_value(sink.acc, neg, -p, pinv);
}

// K primes sharing one packed table: out[k] as glynn_chunk for primes[k].
// kPaired requires every prime below kPairedPrimeLimit.
template <int NR, int G, int K, bool kPaired = false>
__device__ __forceinline__ void glynn_chunk_set(const int32_t* __restrict__ tables, int n, int cbits,
                                                uint64_t base, const int32_t (&primes)[K],
                                                const uint32_t (&pinvs)[K], int32_t (&out)[K]) {
    PrimeSetSink<NR, G, K, kPaired> sink;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        sink.negp[k] = -primes[k];
        sink.pinv[k] = pinvs[k];
        sink.acc[k] = 0;
    }
    const bool neg = gray_walk<NR, G>(tables, n, cbits, base, 0, sink);
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = lane_value(sink.acc[k], neg, sink.negp[k], sink.pinv[k]);
}

// The lane's exact signed sum of terms as L int64 columns (value
// sum_c out[c] 2^(31 c)); ExactSink states which tables each L takes.
template <int NR, int G, int L, int V = 31>
__device__ __forceinline__ void glynn_chunk_exact(const int32_t* __restrict__ tables, int n, int cbits,
                                                  uint64_t base, int64_t (&out)[L]) {
    ExactSink<NR, G, L, V> sink;
#pragma unroll
    for (int c = 0; c < L; ++c) sink.acc[c] = 0;
    const bool neg = gray_walk<NR, G>(tables, n, cbits, base, 0, sink);
#pragma unroll
    for (int c = 0; c < L; ++c) out[c] = neg ? -sink.acc[c] : sink.acc[c];
}

// glynn_chunk_set through ZeroSkipSink, on byte tables (build_byte_tables,
// byte_stride(NR) words per column).  Every lane of the wavefront must
// call it (the queue is shared); lanes with active == false walk but queue
// nothing.  A lane's value is its sum of the absolutely signed terms it
// took, which may come from any lane's chunk, so only the wavefront's total
// is meaningful: it equals the total of glynn_chunk_set's active lanes.
// Each lane takes at most 2^cbits terms (one per full batch of 64 and one
// at the flush), so |acc| < 2^(31+cbits) as there.
template <int NR, int G, int K, bool kHalfTest>
__device__ __forceinline__ void glynn_chunk_set_skip(const int32_t* __restrict__ tables, int n, int cbits,
                                                     uint64_t base, bool active, int32_t* queue,
                                                     const int32_t (&primes)[K], const uint32_t (&pinvs)[K],
                                                     int32_t (&out)[K]) {
    ZeroSkipSink<NR, G, K, kHalfTest> sink;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        sink.negp[k] = -primes[k];
        sink.pinv[k] = pinvs[k];
        sink.acc[k] = 0;
    }
    sink.queue = queue;
    sink.count = 0;
    sink.lane = __lane_id();
    sink.neg = __popcll(base ^ (base >> 1)) & 1;
    sink.active = active;
    sink.active_mask = __ballot(active);
    sink.x1010 = 0xfefefeffu;
    gray_walk_rows<NR, ByteRows<NR>>(tables, n, cbits, base, 0, sink);
    sink.flush();
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = lane_value(sink.acc[k], false, sink.negp[k], sink.pinv[k]);
}


// ---------------------------------------------------------------------------
// Block zero-skipping walk for 0/1 tables (build_block_tables).
//
// The chunk's Gray states are visited as 2^(c-5) outer states (the Gray
// code over columns 5..n-2, one flip per outer step) times all 32 inner
// states y of the first kBlockInnerBits Gray columns: state (u, y) has row
// sums O_i(u) - 2 * (ones of row i among the columns selected by y) and
// sign (-1)^(parity of u's Gray bits + popcount y).  A row of even degree
// is zero at the inner states whose count of ones equals O_i / 2, a set
// that depends only on O_i: the host tabulates it per row as 32-bit masks
// over y, indexed by the row's byte, so one LDS lookup per row (an SDWA
// byte add and a ds_read_b32) tests 32 states.  The states left (every
// tested row nonzero) are queued as entries of a per-wavefront LDS ring:
// the lane's outer row words with the sign, its survivor mask and
// the index of its first state in the wavefront's running count.  Whenever
// 64 states are queued, each lane takes one: it finds its entry by the
// entry's first index (entries starting inside the batch flag their first
// lane; a lane's entry is the number of flags up to it), picks the y of its rank
// in the survivor mask, rebuilds the row sums from the inner table and adds
// the term as ZeroSkipSink does.  Row bytes are doubled sums 2 s + beta
// (beta = 128 even degree, 126 odd), all multiples of 4, so the bytes index
// the mask tables directly.  Entries hold the halved outer words (the sign
// in bit 31), and the inner table the halved, re-biased inner words, so a
// word of a taken state's signed sums is one subtraction and one XOR.

typedef __attribute__((address_space(3))) uint32_t lds_u32;

// Inclusive prefix sum over the wavefront in lane order: DPP row shifts
// inside rows of 16 lanes, then the gfx9 row broadcasts of lanes 15 and 31.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
    return v;
}

// Position of set bit k (counting from 0 at the low end) of m, by halving
// popcounts; always in [0, 31], meaningful for k < popcount(m).
__device__ __forceinline__ uint32_t kth_set_bit(uint32_t m, uint32_t k) {
    uint32_t pos = 0;
#pragma unroll
    for (int level = 0; level < 5; ++level) {
        const int h = 16 >> level;
        const uint32_t c = (uint32_t)__popc((m >> pos) & ((1u << h) - 1u));
        const bool up = k >= c;
        k = up ? k - c : k;
        pos = up ? pos + h : pos;
    }
    return pos;
}

// kMod3: the modulo-3 walk (build_block_tables_mod3): 64-word mask tables
// per row, and the nonzero states, which are rare there (every row must
// avoid a multiple of 3), taken in place by the lanes that hold them.
//
// kExact: the terms go to K = kExactColumns exact limb columns instead of K
// primes' residues: each taken state's product of group products in
// radix-2^31 limbs (LimbProduct, full products of at most eight groups of
// magnitudes below 2^31), each limb added with the term's sign to its int64
// column, for every prime of the job at once.  A lane takes at most one
// state per batch of 64 and at most 2^cbits batches complete per item, so
// |column| < 2^(31 + cbits) per lane, as in ExactSink.
template <int NR, int G, int K, bool kMod3 = false, bool kExact = false>
struct BlockSkipWalk {
    static_assert(G > 0, "prime sets need exact row sums");
    static_assert(!kExact || (!kMod3 && K == kExactColumns && (NR + G - 1) / G <= kExactMaxGroups),
                  "exact block walks take at most eight groups in kExactColumns columns");
    using Rows = ByteRows<NR>;
    using Groups = ByteGroups<NR, G>;
    static constexpr int W = Rows::kStride;
    static constexpr int kGroups = Groups::kGroups;
    static constexpr int kPairs = kGroups / 2;
    static constexpr int kLeaves = kPairs + (kGroups & 1);
    static constexpr uint32_t kRing = kBlockRing;
    int32_t negp[K];
    uint32_t pinv[K];
    int64_t acc[K];
    lds_u32* inner;       // 12 W words: row word k's two inner tables at 12 k (build_block_tables)
    lds_u32* owner;       // 64 words: flags of the lanes where an entry starts
    lds_u32* ring;        // (W + 2) * kRing words
    uint32_t mask_base;   // LDS byte address of row 0's mask table (row i at + 128 i; + 256 i modulo 3)
    int test_words;       // rows tested / 4 (wave-uniform)
    uint32_t ehead, etail, thead, ttail;  // entries and states queued / taken (wave-uniform)
    uint32_t lane;
    bool active;

    // Mask entry of row `row` (byte B of its row word w): the byte plus the
    // table base is the entry's LDS byte address (one SDWA add, which the
    // compiler forms from the shift and mask) and the row's table offset
    // goes into the ds_read_b32 offset field.
    template <int B>
    __device__ __forceinline__ uint32_t lookup(uint32_t w, int row) const {
        const uint32_t addr = ((w >> (8 * B)) & 0xffu) + mask_base;
        return ((const lds_u32*)(uintptr_t)addr)[row * (kMod3 ? 64 : 32)];
    }

    // Adds the term of one state from its signed row-sum bytes w and its
    // sign (1: negative): the group products, the pair products and each
    // prime's Montgomery tree, as ZeroSkipSink does.
    __device__ __forceinline__ void add_term(const uint32_t (&w)[W], uint32_t sign) {
        int32_t g[kGroups];
        const Groups q{w};
        q.groups(g, std::make_integer_sequence<int, kGroups>{});
        if constexpr (kExact) {
            uint32_t m[kGroups];
#pragma unroll
            for (int k = 0; k < kGroups; ++k) m[k] = (uint32_t)(g[k] < 0 ? -g[k] : g[k]);
            // The term's sign: bit 31 of the XOR of the group products and
            // of the state's sign, folded three at a time.
            uint32_t s = sign << 31;
#pragma unroll
            for (int k = 0; k < kGroups; k += 2)
                s = k + 1 < kGroups ? xor3_fast(s, (uint32_t)g[k], (uint32_t)g[k + 1]) : s ^ (uint32_t)g[k];
            const int32_t e = ((int32_t)s >> 31) | 1;
            using P = LimbProduct<0, kGroups, kExactColumns>;
            uint32_t z[P::kLimbs];
            P::run(m, z);
#pragma unroll
            for (int c = 0; c < P::kLimbs; ++c) acc[c] = mad_exact(acc[c], (int32_t)z[c], e);
            return;
        }
        if (sign) g[0] = -g[0];
        int64_t P[kPairs > 0 ? kPairs : 1];
#pragma unroll
        for (int j = 0; j < kPairs; ++j) P[j] = (int64_t)g[2 * j] * (int64_t)g[2 * j + 1];
#pragma unroll
        for (int p = 0; p < K; ++p) {
            int32_t v[kLeaves];
#pragma unroll
            for (int j = 0; j < kPairs; ++j) v[j] = redc31(P[j], negp[p], pinv[p]);
            if constexpr (kGroups & 1) v[kLeaves - 1] = g[kGroups - 1];
#pragma unroll
            for (int wd = 1; wd < kLeaves; wd *= 2) {
#pragma unroll
                for (int i = 0; i + wd < kLeaves; i += 2 * wd) v[i] = mont31(v[i], v[i + wd], negp[p], pinv[p]);
            }
            acc[p] = mac_sign<1>(acc[p], v[0]);
        }
    }

    // Bit y set: some tested row is zero at inner state y.  Rows are tested
    // eight at a time (two row words), whose eight loads are in flight
    // together; test_words is even or the table's last word.
    __device__ __forceinline__ uint32_t zero_mask(const Rows& rs) const {
        uint32_t m = 0;
#pragma unroll
        for (int k = 0; k < W; k += 2) {
            if (k < test_words) {
                uint32_t a[8];
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    if (k + b / 4 < W) {
                        const uint32_t w = rs.r[k + b / 4];
                        switch (b % 4) {
                            case 0: a[b] = lookup<0>(w, 4 * k + b); break;
                            case 1: a[b] = lookup<1>(w, 4 * k + b); break;
                            case 2: a[b] = lookup<2>(w, 4 * k + b); break;
                            default: a[b] = lookup<3>(w, 4 * k + b); break;
                        }
                    } else {
                        a[b] = 0u;
                    }
                }
#pragma unroll
                for (int b = 0; b < 8; b += 2) m = __builtin_amdgcn_bitop3_b32(m, a[b], a[b + 1], 0xfe);
            }
        }
        return m;
    }

    // Takes `cnt` (<= 64) queued states, one per lane from lane 0.
    __device__ __forceinline__ void consume(uint32_t cnt) {
        __builtin_amdgcn_wave_barrier();
        owner[lane] = 0u;
        __builtin_amdgcn_wave_barrier();
        // Every entry holds at least one state, so the entries that start
        // inside the batch are among the 64 after the oldest live one
        // (etail, which holds state ttail and is every lane's default).
        const uint32_t live = ehead - etail;
        if (lane + 1u < live) {
            const uint32_t slot = (etail + lane + 1u) & (kRing - 1);
            const uint32_t rel = ring[(W + 1) * kRing + slot] - ttail;
            if (rel < 64u) owner[rel] = 1u;
        }
        __builtin_amdgcn_wave_barrier();
        // Entries start in order, so a lane's entry (relative to etail) is
        // the number of entries that start at or before it.
        const uint32_t flag = owner[lane];
        const uint64_t starts = __ballot(flag != 0u);
        const uint32_t rel_e =
            __builtin_amdgcn_mbcnt_hi((uint32_t)(starts >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)starts, 0u)) +
            flag;
        const uint32_t slot = (etail + rel_e) & (kRing - 1);
        const uint32_t k = ttail + lane - ring[(W + 1) * kRing + slot];
        const uint32_t mask = ring[W * kRing + slot];
        const uint32_t y = kth_set_bit(mask, k);
        uint32_t w[W];
        uint32_t sign = 0;
        // The halved outer word plus the two inner words has the bytes
        // s + 128 (build_block_tables); flipping bit 7 gives the signed sums.
        const uint32_t ya = y & 3u, yb = 4u + (y >> 2);
#pragma unroll
        for (int kk = 0; kk < W; ++kk) {
            uint32_t o = ring[kk * kRing + slot];
            if (kk == 0) {
                sign = o >> 31;
                o &= 0x7fffffffu;
            }
            w[kk] = (o + inner[kk * 12 + ya] + inner[kk * 12 + yb]) ^ 0x80808080u;
        }
        sign ^= (uint32_t)__popc(y) & 1u;
        if (lane < cnt) add_term(w, sign);
        if (cnt == 64u) {
            // The next state to take lies in lane 63's entry unless lane 63
            // took that entry's last state.
            const uint32_t next = rel_e + ((k + 1u == (uint32_t)__popc(mask)) ? 1u : 0u);
            etail += (uint32_t)__builtin_amdgcn_readlane((int)next, 63);
        }
        ttail += cnt;
    }

    // One outer state: tests its 32 inner states and queues the survivors
    // (modulo 3: takes them in place).
    __device__ __forceinline__ void block(const Rows& rs, uint32_t par) {
        const uint32_t surv = ~zero_mask(rs);
        if constexpr (kMod3) {
            uint32_t left = active ? surv : 0u;
            while (__ballot(left != 0u)) {
                if (left != 0u) {
                    const uint32_t y = (uint32_t)__builtin_ctz(left);
                    left &= left - 1u;
                    const uint32_t ya = y & 3u, yb = 4u + (y >> 2);
                    uint32_t w[W];
#pragma unroll
                    for (int kk = 0; kk < W; ++kk)
                        w[kk] = ((rs.r[kk] >> 1) + inner[kk * 12 + ya] + inner[kk * 12 + yb]) ^ 0x80808080u;
                    add_term(w, par ^ ((uint32_t)__popc(y) & 1u));
                }
            }
            return;
        }
        const bool has = active && surv != 0u;
        const uint64_t bal = __ballot(has);
        if (bal == 0) return;
        const uint32_t c = has ? (uint32_t)__popc(surv) : 0u;
        const uint32_t incl = wave_inclusive_sum(c);
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        if (has) {
            const uint32_t rank =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            const uint32_t slot = (ehead + rank) & (kRing - 1);
            // Halved outer words (every byte is even); word 0 carries the
            // outer parity in bit 31 (free: a halved top byte is below 128).
#pragma unroll
            for (int kk = 0; kk < W; ++kk)
                ring[kk * kRing + slot] = kk == 0 ? __builtin_amdgcn_alignbit(par, rs.r[0], 1) : rs.r[kk] >> 1;
            ring[W * kRing + slot] = surv;
            ring[(W + 1) * kRing + slot] = thead + incl - c;
        }
        ehead += (uint32_t)__popcll(bal);
        thead += total;
        while (thead - ttail >= 64u) consume(64u);
    }
};

// glynn_chunk_set_skip's block form on block tables (build_block_tables):
// walks the lane's 2^cbits states (cbits >= kBlockMinChunkBits) into
// walk.acc (the K primes' Montgomery sums, or the exact limb columns; only
// the wavefront's total is meaningful, as there).  `region` is the
// wavefront's LDS region (block_region_words(n, rows) words); with `reload`
// the mask and inner tables are copied into it first (a wavefront keeps
// them while it stays on one job).  Every lane of the wavefront must call
// it.
template <int NR, int G, int K, bool kMod3, bool kExact>
__device__ __forceinline__ void block_walk_chunk(BlockSkipWalk<NR, G, K, kMod3, kExact>& walk,
                                                 const int32_t* __restrict__ tables, int n, int cbits, uint64_t base,
                                                 bool active, int32_t* region, int rows, bool reload) {
    using Walk = BlockSkipWalk<NR, G, K, kMod3, kExact>;
    constexpr int W = Walk::W;
#pragma unroll
    for (int k = 0; k < K; ++k) walk.acc[k] = 0;
    const uint32_t lane = __lane_id();
    lds_u32* const lds = (lds_u32*)region;
    // Mask tables (rows + 1 overlapping tables of 32 words, or rows of 64
    // modulo 3), then the inner tables, the owner flags and the ring.
    const int mask_words = kMod3 ? rows * 64 : (rows + 1) * 32;
    walk.inner = lds + mask_words;
    walk.owner = walk.inner + 12 * W;
    walk.ring = walk.owner + 64;
    walk.mask_base = (uint32_t)(uintptr_t)lds;
    walk.test_words = rows / 4;
    walk.ehead = walk.etail = walk.thead = walk.ttail = 0;
    walk.lane = lane;
    walk.active = active;
    const int32_t* __restrict__ D = tables + W;
    if (reload) {
        const int32_t* __restrict__ src = tables + (size_t)n * W;
        for (int i = (int)lane; i < 12 * W; i += 64) walk.inner[i] = (uint32_t)src[i];
        src += 12 * W;
        for (int i = (int)lane; i < mask_words; i += 64) lds[i] = (uint32_t)src[i];
    }
    __builtin_amdgcn_wave_barrier();

    // Outer row sums of the chunk's first state: the Gray bits of base from
    // column kBlockInnerBits on (bits below cbits-1 are zero).
    const uint64_t gray = base ^ (base >> 1);
    ByteRows<NR> rs;
    rs.load(tables);
    for (int j = cbits - 1; j < n - 1; ++j) rs.sub_masked(D + j * W, -(int32_t)((gray >> j) & 1u), 0);
    uint32_t par = (uint32_t)__popcll(gray) & 1u;
    const uint64_t ub = base >> kBlockInnerBits;
    const int ubits = cbits - kBlockInnerBits;
    const uint32_t U = 1u << ubits;
    walk.block(rs, par);
    for (uint32_t u = 1; u < U; ++u) {
        const int jj = __builtin_ctz(u);  // uniform
        const int32_t* col = D + (kBlockInnerBits + jj) * W;
        // The new Gray bit is 1 ^ bit jj+1 of (ub + u): bit jj+1 of u except
        // at the chunk's middle outer state, where it is the lane's own.
        if (jj + 1 < ubits) {
            if ((u >> (jj + 1)) & 1u) {
                rs.template flip_mem<false>(col, 0);
            } else {
                rs.template flip_mem<true>(col, 0);
            }
        } else {
            rs.flip_lane(col, !(((ub + u) >> (jj + 1)) & 1u), 0);
        }
        par ^= 1u;
        walk.block(rs, par);
    }
    if constexpr (!kMod3) {
        const uint32_t pend = walk.thead - walk.ttail;
        if (pend) walk.consume(pend);
    }
}

// The residues of the K primes in out (each lane's Montgomery sum reduced).
template <int NR, int G, int K, bool kMod3 = false>
__device__ __forceinline__ void glynn_chunk_set_blocks(const int32_t* __restrict__ tables, int n, int cbits,
                                                       uint64_t base, bool active, int32_t* region, int rows,
                                                       bool reload, const int32_t (&primes)[K],
                                                       const uint32_t (&pinvs)[K], int32_t (&out)[K]) {
    BlockSkipWalk<NR, G, K, kMod3> walk;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        walk.negp[k] = -primes[k];
        walk.pinv[k] = pinvs[k];
    }
    block_walk_chunk(walk, tables, n, cbits, base, active, region, rows, reload);
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = lane_value(walk.acc[k], false, walk.negp[k], walk.pinv[k]);
}

// The same walk with exact accumulation: the lane's kExactColumns limb
// columns of the Glynn sum over the states it takes (only the wavefront's
// total is meaningful).
template <int NR, int G>
__device__ __forceinline__ void glynn_chunk_blocks_exact(const int32_t* __restrict__ tables, int n, int cbits,
                                                         uint64_t base, bool active, int32_t* region, int rows,
                                                         bool reload, int64_t (&cols)[kExactColumns]) {
    BlockSkipWalk<NR, G, kExactColumns, false, true> walk;
    block_walk_chunk(walk, tables, n, cbits, base, active, region, rows, reload);
#pragma unroll
    for (int k = 0; k < kExactColumns; ++k) cols[k] = walk.acc[k];
}

}  // namespace phip
