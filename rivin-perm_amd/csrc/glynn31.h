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
// t = 2^(c-1), which reads bit c of the lane's base.  A launch covers every
// prime at once: blockIdx.y selects the prime slice.
//
// Row-sum representations (template parameter G):
//   G = 0     row sums are residues in [0, p); a flip adds or subtracts
//             d = 2a mod p and reduces with one unsigned min.
//   G = 2..4  "small" matrices whose centred residues have maximum absolute
//             row sum S with S^G < 2^31 and S^(G-1) < 2^23: row sums are the
//             exact integers (one add per row, no reduction) and each group
//             of G row sums is multiplied exactly with 24-bit multiplies
//             before one Montgomery step, so a state costs ceil(n/G)-1
//             Montgomery products instead of n-1.
// Products are combined with the signed Montgomery multiply in a balanced
// tree; the accumulated sum therefore carries R^-(groups-1), which the
// finalisation step removes together with 2^-(n-1).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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

// Hides a pointer from loop-invariant code motion so each flip issues its
// own scalar loads instead of pinning every column in SGPRs.
__device__ __forceinline__ const int32_t* opaque_ptr(const int32_t* p) {
    asm volatile("" : "+s"(p));
    return p;
}

template <int NR, int G>
struct GlynnRows {
    static constexpr bool kWide = (G == 0);
    static constexpr int kGroups = kWide ? NR : (NR + G - 1) / G;

    // Product of the row sums times R^-(kGroups-1); |result| < 2^31.
    __device__ __forceinline__ static int32_t term(const int32_t (&rs)[NR], int32_t negp,
                                                   uint32_t pinv) {
        int32_t v[kGroups];
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            if constexpr (kWide) {
                v[g] = rs[g];
            } else {
                int32_t gp = rs[g * G];
#pragma unroll
                for (int e = 1; e < G; ++e) {
                    if (g * G + e < NR) gp = mul_i24(gp, rs[g * G + e]);
                }
                v[g] = gp;
            }
        }
#pragma unroll
        for (int w = 1; w < kGroups; w *= 2) {
#pragma unroll
            for (int i = 0; i + w < kGroups; i += 2 * w) {
                v[i] = mont31(v[i], v[i + w], negp, pinv);
            }
        }
        return v[0];
    }

    __device__ __forceinline__ static int32_t add_row(int32_t r, int32_t d, int32_t p) {
        if constexpr (kWide) {
            uint32_t x = (uint32_t)r + (uint32_t)d;
            uint32_t y = x - (uint32_t)p;
            return (int32_t)(x < y ? x : y);
        } else {
            return r + d;
        }
    }

    __device__ __forceinline__ static int32_t sub_row(int32_t r, int32_t d, int32_t p) {
        if constexpr (kWide) {
            uint32_t x = (uint32_t)r - (uint32_t)d;
            uint32_t y = x + (uint32_t)p;
            return (int32_t)(x < y ? x : y);
        } else {
            return r - d;
        }
    }

    template <bool SUB>
    __device__ __forceinline__ static void flip(int32_t (&rs)[NR], const int32_t (&d)[NR], int32_t p) {
#pragma unroll
        for (int i = 0; i < NR; ++i) rs[i] = SUB ? sub_row(rs[i], d[i], p) : add_row(rs[i], d[i], p);
    }

    template <bool SUB>
    __device__ __forceinline__ static void flip_mem(int32_t (&rs)[NR], const int32_t* __restrict__ col,
                                                    int32_t p) {
        const int32_t* c = opaque_ptr(col);
#pragma unroll
        for (int i = 0; i < NR; ++i) rs[i] = SUB ? sub_row(rs[i], c[i], p) : add_row(rs[i], c[i], p);
    }

    // Flip whose direction may differ between lanes.
    __device__ __forceinline__ static void flip_lane(int32_t (&rs)[NR], const int32_t* __restrict__ col,
                                                     bool sub, int32_t p) {
        const int32_t* c = opaque_ptr(col);
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const int32_t d = c[i];
            if constexpr (kWide) {
                const int32_t e = sub ? p - d : d;  // p - 0 = p still reduces to rs
                rs[i] = add_row(rs[i], e, p);
            } else {
                rs[i] += sub ? -d : d;
            }
        }
    }
};

// One lane's chunk: the 2^cbits Gray states starting at `base` (a multiple
// of 2^cbits).  rs0 holds the NR row sums of state 0 (all column signs +1:
// exact centred sums when G > 0, residues when G = 0; padding rows hold 1),
// D the flip deltas of the n-1 Gray columns, column-major with stride NR
// (2a_ij centred, or 2a_ij mod p; padding rows hold 0).  Returns the lane's
// signed sum of terms reduced into [0, p); it still carries the factor
// R^-(groups-1) and lacks 2^-(n-1).
template <int NR, int G>
__device__ __forceinline__ uint64_t glynn_chunk(const int32_t* __restrict__ rs0,
                                                const int32_t* __restrict__ D, int n, int cbits,
                                                uint64_t base, int32_t p, uint32_t pinv) {
    using Rows = GlynnRows<NR, G>;
    const int32_t negp = -p;
    const uint64_t gray = base ^ (base >> 1);

    // Row sums of the chunk's first state: subtract 2a for every set Gray
    // bit (perm_c_omp.c:582-600).  Bits below c-1 of gray(base) are zero.
    int32_t rs[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) rs[i] = rs0[i];
    for (int j = cbits - 1; j < n - 1; ++j) {
        const int32_t mask = -(int32_t)((gray >> j) & 1u);
        const int32_t* __restrict__ col = D + j * NR;
#pragma unroll
        for (int i = 0; i < NR; ++i) rs[i] = Rows::sub_row(rs[i], col[i] & mask, p);
    }

    // Column 0 flips every other state; keep it resident.
    int32_t c0[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) c0[i] = D[i];
    const int32_t* __restrict__ d1 = D + NR;
    const int32_t* __restrict__ d2 = D + 2 * NR;

    // Terms at even local index carry the chunk's starting sign, odd ones
    // the opposite sign.
    int64_t acc_e = Rows::term(rs, negp, pinv);
    int64_t acc_o = 0;
    const uint32_t T = 1u << cbits;
    for (uint32_t t0 = 0; t0 < T; t0 += kGlynnUnroll) {
        // Directions for s = 1..15 follow from bit j+1 of t0+s, which lies
        // inside the unrolled block except for s = 8 (bit 4 of base+t0).
        Rows::template flip<true>(rs, c0, p);         acc_o += Rows::term(rs, negp, pinv);  // s=1
        Rows::template flip_mem<true>(rs, d1, p);     acc_e += Rows::term(rs, negp, pinv);  // s=2
        Rows::template flip<false>(rs, c0, p);        acc_o += Rows::term(rs, negp, pinv);  // s=3
        Rows::template flip_mem<true>(rs, d2, p);     acc_e += Rows::term(rs, negp, pinv);  // s=4
        Rows::template flip<true>(rs, c0, p);         acc_o += Rows::term(rs, negp, pinv);  // s=5
        Rows::template flip_mem<false>(rs, d1, p);    acc_e += Rows::term(rs, negp, pinv);  // s=6
        Rows::template flip<false>(rs, c0, p);        acc_o += Rows::term(rs, negp, pinv);  // s=7
        {
            // s=8 flips column 3; its new Gray bit is 1 ^ bit 4 of (base+t).
            const uint64_t t = base + t0 + 8;
            Rows::flip_lane(rs, D + 3 * NR, !((t >> 4) & 1u), p);
            acc_e += Rows::term(rs, negp, pinv);
        }
        Rows::template flip<true>(rs, c0, p);         acc_o += Rows::term(rs, negp, pinv);  // s=9
        Rows::template flip_mem<true>(rs, d1, p);     acc_e += Rows::term(rs, negp, pinv);  // s=10
        Rows::template flip<false>(rs, c0, p);        acc_o += Rows::term(rs, negp, pinv);  // s=11
        Rows::template flip_mem<false>(rs, d2, p);    acc_e += Rows::term(rs, negp, pinv);  // s=12
        Rows::template flip<true>(rs, c0, p);         acc_o += Rows::term(rs, negp, pinv);  // s=13
        Rows::template flip_mem<false>(rs, d1, p);    acc_e += Rows::term(rs, negp, pinv);  // s=14
        Rows::template flip<false>(rs, c0, p);        acc_o += Rows::term(rs, negp, pinv);  // s=15
        const uint32_t tn = t0 + kGlynnUnroll;
        if (tn < T) {
            const int j = __builtin_ctz(tn);  // 4 <= j <= cbits-1, uniform
            const uint64_t t = base + tn;
            Rows::flip_lane(rs, D + j * NR, !((t >> (j + 1)) & 1u), p);
            acc_e += Rows::term(rs, negp, pinv);
        }
    }

    int64_t acc = acc_e - acc_o;
    if (__popcll(gray) & 1) acc = -acc;
    int64_t r = acc % (int64_t)p;
    if (r < 0) r += p;
    return (uint64_t)r;
}

// Per-prime slice layout (all int32):
//   rs0    [k][NR]       row sums of state 0 (see glynn_chunk)
//   deltas [k][n-1][NR]  flip deltas (see glynn_chunk)
//   partials[k][gridDim.x] receives each block's sum of lane residues.
template <int NR, int G>
__global__ __launch_bounds__(kGlynnBlock) void glynn31_kernel(
    const int32_t* __restrict__ rs0, const int32_t* __restrict__ deltas,
    const int32_t* __restrict__ primes, const uint32_t* __restrict__ pinvs, int n, int cbits,
    uint64_t lanes_per_prime, uint64_t* __restrict__ partials) {
    const int q = blockIdx.y;
    const uint64_t lane = (uint64_t)blockIdx.x * kGlynnBlock + threadIdx.x;
    const bool active = lane < lanes_per_prime;
    const uint64_t r = glynn_chunk<NR, G>(rs0 + (size_t)q * NR, deltas + (size_t)q * (n - 1) * NR, n, cbits,
                                          (active ? lane : 0) << cbits, primes[q], pinvs[q]);
    const uint64_t block_sum = block_reduce_sum_u64<kGlynnBlock>(active ? r : 0u);
    if (threadIdx.x == 0) partials[(size_t)q * gridDim.x + blockIdx.x] = block_sum;
}

}  // namespace phip
