// Bit-sliced permanent modulo 3 (gfx950).
//
// Elements of F_3 are stored in the reference's "bipedal" encoding
// (perm_mod3_omp.c:1-37): a pair of bits (mag, sgn) per element, with
// 0 -> (0,0), 1 -> (1,0), -1 -> (1,1).  A whole column of n < 64 entries is
// a pair of 64-bit words, and a vector add or subtract is a handful of
// bitwise operations on those words.  The product of n row sums is nonzero
// iff every mag bit is set, and then equals (-1)^popcount(sgn).
//
// The reference enumerates all 2^n Ryser subsets (perm_mod3_omp.c:54-156).
// Here the fixed-sign Glynn identity is used instead, which is valid in any
// ring where 2 is invertible and halves the work to 2^(n-1) states:
//     per(A) = 2^-(n-1) sum_{d, d_n = +1} (prod_k d_k) prod_i (sum_j d_j a_ij)
// Over F_3, 2 = -1, so flipping column j from + to - adds a_j (rs - 2a = rs + a)
// and flipping it back subtracts a_j, and 2^-(n-1) = (-1)^(n-1).
//
// Each lane owns 2^c consecutive Gray states; the flipped column is
// wave-uniform (scalar loads), the row-sum words live in VGPRs, and the
// nonzero test and sign are v_cmp and v_bcnt.  Lane counts are summed with
// DPP/swizzle and LDS, block counts by a second pass (no atomics).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "device_context.h"
#include "phip_reduce.h"
#include "residues.h"

namespace phip {

namespace {

constexpr int kMod3Block = 256;

// Packed F_3 vector; W = 1 for n <= 32, W = 2 otherwise.
template <int W>
struct Vec3 {
    uint32_t m[W];
    uint32_t s[W];
};

// (m1,s1) + (m2,s2)  (perm_mod3_omp.c:21-28)
template <int W>
__device__ __forceinline__ void f3_add(Vec3<W>& a, const uint32_t (&m2)[W], const uint32_t (&s2)[W]) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint32_t t = a.m[w] ^ a.s[w] ^ s2[w];
        const uint32_t u = m2[w] & t;
        const uint32_t mag = u | (a.m[w] ^ m2[w]);
        a.s[w] = u ^ a.s[w];
        a.m[w] = mag;
    }
}

// (m1,s1) - (m2,s2)  (perm_mod3_omp.c:30-37)
template <int W>
__device__ __forceinline__ void f3_sub(Vec3<W>& a, const uint32_t (&m2)[W], const uint32_t (&s2)[W]) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint32_t t = a.s[w] ^ s2[w];
        const uint32_t u = a.m[w] & t;
        const uint32_t mag = u | (a.m[w] ^ m2[w]);
        a.s[w] = u ^ (m2[w] ^ s2[w]);
        a.m[w] = mag;
    }
}

// +1, -1 or 0 for the product of the row sums.
template <int W>
__device__ __forceinline__ int f3_product(const Vec3<W>& a, const uint32_t (&mask)[W]) {
    bool nz = true;
    uint32_t bits = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        nz = nz && (a.m[w] == mask[w]);
        bits += __popc(a.s[w] & mask[w]);
    }
    const int sign = 1 - 2 * (int)(bits & 1u);
    return nz ? sign : 0;
}

template <int W>
__device__ __forceinline__ void load_col(const uint32_t* __restrict__ cols, int j, uint32_t (&m)[W],
                                         uint32_t (&s)[W]) {
    const uint32_t* c = cols + (size_t)j * 2 * W;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        m[w] = c[w];
        s[w] = c[W + w];
    }
}

// Column words through the constant address space: the flipped column is
// the same for every lane, so its words come in by scalar loads; the empty
// asm keeps the compiler from hoisting every column into registers.
typedef const __attribute__((address_space(4))) uint32_t* ConstWords;
__device__ __forceinline__ ConstWords const_words(const uint32_t* p) {
    asm volatile("" : "+s"(p));
    return (ConstWords)p;
}

template <int W>
__device__ __forceinline__ void load_col_uniform(const uint32_t* __restrict__ cols, int j, uint32_t (&m)[W],
                                                 uint32_t (&s)[W]) {
    const ConstWords c = const_words(cols + (size_t)j * 2 * W);
#pragma unroll
    for (int w = 0; w < W; ++w) {
        m[w] = c[w];
        s[w] = c[W + w];
    }
}

template <int W, bool ADD>
__device__ __forceinline__ void f3_flip(Vec3<W>& a, const uint32_t (&m)[W], const uint32_t (&s)[W]) {
    if constexpr (ADD) {
        f3_add<W>(a, m, s);
    } else {
        f3_sub<W>(a, m, s);
    }
}

// A flip whose direction differs between lanes: both results, then a select.
template <int W>
__device__ __forceinline__ void f3_flip_lane(Vec3<W>& a, const uint32_t (&m)[W], const uint32_t (&s)[W], bool add) {
    Vec3<W> x = a, y = a;
    f3_add<W>(x, m, s);
    f3_sub<W>(y, m, s);
#pragma unroll
    for (int w = 0; w < W; ++w) {
        a.m[w] = add ? x.m[w] : y.m[w];
        a.s[w] = add ? x.s[w] : y.s[w];
    }
}

// f3_product for row sums whose bits beyond the order are zero (sign bits
// lie inside the magnitude bits, so no mask is needed for the parity).
template <int W>
__device__ __forceinline__ int f3_term(const Vec3<W>& a, const uint32_t (&mask)[W]) {
    bool nz = true;
    uint32_t bits = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        nz = nz && (a.m[w] == mask[w]);
        bits += __popc(a.s[w]);
    }
    return nz ? 1 - 2 * (int)(bits & 1u) : 0;
}

// cols[j] = {mag words..., sgn words...}; rs0 = column sum with all signs +.
template <int W>
__global__ __launch_bounds__(kMod3Block) void mod3_glynn_kernel(const uint32_t* __restrict__ cols,
                                                                const uint32_t* __restrict__ rs0, int n,
                                                                int cbits, uint64_t lanes,
                                                                int64_t* __restrict__ partials) {
    const uint64_t lane = (uint64_t)blockIdx.x * kMod3Block + threadIdx.x;
    const bool active = lane < lanes;
    const uint64_t base = (active ? lane : 0) << cbits;
    const uint64_t gray = base ^ (base >> 1);
    uint32_t mask[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const int lo = 32 * w;
        mask[w] = (n >= lo + 32) ? 0xffffffffu : (n <= lo ? 0u : ((1u << (n - lo)) - 1u));
    }
    Vec3<W> rs;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        rs.m[w] = rs0[w];
        rs.s[w] = rs0[W + w];
    }
    // Columns whose Gray bit is set carry sign -1: rs += a_j.
    for (int j = (cbits > 0 ? cbits - 1 : 0); j < n - 1; ++j) {
        if ((gray >> j) & 1u) {
            uint32_t m[W], s[W];
            load_col<W>(cols, j, m, s);
            f3_add<W>(rs, m, s);
        }
    }
    const uint64_t T = 1ull << cbits;
    // Terms of even local index add, odd ones subtract (the chunk's first
    // state's sign is applied at the end).  |total| <= 2^cbits may pass
    // 2^31 at large orders, so each step's int sum is folded into int64.
    int64_t total = f3_term<W>(rs, mask);
    if (cbits >= 4) {
        // Sixteen states per step with the flip directions of the Gray code
        // fixed at compile time: moving to local index t flips column
        // j = ctz(t) to the new Gray bit 1 ^ bit j+1 of t (bits below cbits
        // of base + t are those of t).  Column 0 (every other state) and
        // columns 1 and 2 stay in registers; column 3 (t = 8 mod 16) and the
        // column at each multiple of 16 are loaded, with the direction read
        // from base + t, which may differ between lanes.
        uint32_t c0m[W], c0s[W], c1m[W], c1s[W], c2m[W], c2s[W];
        load_col_uniform<W>(cols, 0, c0m, c0s);
        load_col_uniform<W>(cols, 1, c1m, c1s);
        load_col_uniform<W>(cols, 2, c2m, c2s);
        for (uint64_t t0 = 0; t0 < T; t0 += 16) {
            int acc = 0;
            f3_flip<W, true>(rs, c0m, c0s);  acc -= f3_term<W>(rs, mask);   // t0+1
            f3_flip<W, true>(rs, c1m, c1s);  acc += f3_term<W>(rs, mask);   // t0+2
            f3_flip<W, false>(rs, c0m, c0s); acc -= f3_term<W>(rs, mask);   // t0+3
            f3_flip<W, true>(rs, c2m, c2s);  acc += f3_term<W>(rs, mask);   // t0+4
            f3_flip<W, true>(rs, c0m, c0s);  acc -= f3_term<W>(rs, mask);   // t0+5
            f3_flip<W, false>(rs, c1m, c1s); acc += f3_term<W>(rs, mask);   // t0+6
            f3_flip<W, false>(rs, c0m, c0s); acc -= f3_term<W>(rs, mask);   // t0+7
            {
                uint32_t m[W], s[W];
                load_col_uniform<W>(cols, 3, m, s);
                f3_flip_lane<W>(rs, m, s, !(((base + t0 + 8) >> 4) & 1u));
                acc += f3_term<W>(rs, mask);                                  // t0+8
            }
            f3_flip<W, true>(rs, c0m, c0s);  acc -= f3_term<W>(rs, mask);   // t0+9
            f3_flip<W, true>(rs, c1m, c1s);  acc += f3_term<W>(rs, mask);   // t0+10
            f3_flip<W, false>(rs, c0m, c0s); acc -= f3_term<W>(rs, mask);   // t0+11
            f3_flip<W, false>(rs, c2m, c2s); acc += f3_term<W>(rs, mask);   // t0+12
            f3_flip<W, true>(rs, c0m, c0s);  acc -= f3_term<W>(rs, mask);   // t0+13
            f3_flip<W, false>(rs, c1m, c1s); acc += f3_term<W>(rs, mask);   // t0+14
            f3_flip<W, false>(rs, c0m, c0s); acc -= f3_term<W>(rs, mask);   // t0+15
            const uint64_t tn = t0 + 16;
            if (tn < T) {
                const int j = __builtin_ctzll(tn);  // 4 <= j < cbits, the same in every lane
                uint32_t m[W], s[W];
                load_col_uniform<W>(cols, j, m, s);
                f3_flip_lane<W>(rs, m, s, !(((base + tn) >> (j + 1)) & 1u));
                acc += f3_term<W>(rs, mask);                                  // tn (even)
            }
            total += acc;
        }
    } else {
        for (uint64_t t = 1; t < T; ++t) {
            const int j = __builtin_ctzll(t);
            const uint64_t g = base + t;
            const bool now_negative = ((g ^ (g >> 1)) >> j) & 1u;
            uint32_t m[W], s[W];
            load_col<W>(cols, j, m, s);
            if (now_negative) {
                f3_add<W>(rs, m, s);
            } else {
                f3_sub<W>(rs, m, s);
            }
            const int v = f3_term<W>(rs, mask);
            total += (t & 1) ? -v : v;
        }
    }
    int64_t signed_acc = total;
    if (__popcll(gray) & 1) signed_acc = -signed_acc;
    // Bias into [0, 2^62) so the unsigned block sum is exact, undone on the host.
    const uint64_t biased = active ? (uint64_t)(signed_acc + (int64_t)T) : 0u;
    const uint64_t s = block_reduce_sum_u64<kMod3Block>(biased);
    if (threadIdx.x == 0) partials[blockIdx.x] = (int64_t)s;
}

}  // namespace

// Residue mod 3 from the column encoding; returns 0..2 or -1.
int mod3_from_columns(DeviceContext* ctx, const uint64_t* mags, const uint64_t* sgns, int n) {
    if (n < 0 || n >= 64) {
        set_error("matrix size must be less than 64");
        return -1;
    }
    if (n == 0) return 1;
    if (!mags || !sgns) {
        set_error("column pointers are null");
        return -1;
    }
    const uint64_t rowmask = (n == 64) ? ~0ull : ((1ull << n) - 1);
    if (n == 1) {
        if (!(mags[0] & 1u)) return 0;
        return (sgns[0] & 1u) ? 2 : 1;
    }
    const int W = n <= 32 ? 1 : 2;
    // Column words and the all-plus row sum (reference bipedal add).
    std::vector<uint32_t> cols((size_t)n * 2 * W);
    uint64_t sm = 0, ss = 0;
    for (int j = 0; j < n; ++j) {
        const uint64_t m = mags[j] & rowmask;
        const uint64_t s = sgns[j] & m;
        for (int w = 0; w < W; ++w) {
            cols[(size_t)j * 2 * W + w] = (uint32_t)(m >> (32 * w));
            cols[(size_t)j * 2 * W + W + w] = (uint32_t)(s >> (32 * w));
        }
        const uint64_t t = sm ^ ss ^ s;
        const uint64_t u = m & t;
        const uint64_t mag = u | (sm ^ m);
        ss = u ^ ss;
        sm = mag;
    }
    std::vector<uint32_t> rs0(2 * W);
    for (int w = 0; w < W; ++w) {
        rs0[w] = (uint32_t)(sm >> (32 * w));
        rs0[W + w] = (uint32_t)(ss >> (32 * w));
    }
    const int total = n - 1;  // Gray bits
    int cbits = std::max(0, total - 18);
    if (total >= 4) cbits = std::max(cbits, 4);
    cbits = std::min(cbits, total);
    const uint64_t lanes = 1ull << (total - cbits);
    const uint64_t gx = (lanes + kMod3Block - 1) / kMod3Block;

    hipStream_t st = ctx->stream();
    auto* d_cols = (uint32_t*)ctx->buffer(0, cols.size() * 4);
    auto* d_rs0 = (uint32_t*)ctx->buffer(1, rs0.size() * 4);
    auto* d_part = (int64_t*)ctx->buffer(4, gx * 8);
    auto* h_part = (int64_t*)ctx->host_buffer(3, gx * 8);
    if (!d_cols || !d_rs0 || !d_part || !h_part) return -1;
    PHIP_HIP_TRY(hipMemcpyAsync(d_cols, cols.data(), cols.size() * 4, hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d_rs0, rs0.data(), rs0.size() * 4, hipMemcpyHostToDevice, st));
    if (W == 1) {
        hipLaunchKernelGGL(mod3_glynn_kernel<1>, dim3((unsigned)gx), dim3(kMod3Block), 0, st, d_cols, d_rs0, n,
                           cbits, lanes, d_part);
    } else {
        hipLaunchKernelGGL(mod3_glynn_kernel<2>, dim3((unsigned)gx), dim3(kMod3Block), 0, st, d_cols, d_rs0, n,
                           cbits, lanes, d_part);
    }
    PHIP_HIP_TRY(hipGetLastError());
    PHIP_HIP_TRY(hipMemcpyAsync(h_part, d_part, gx * 8, hipMemcpyDeviceToHost, st));
    PHIP_HIP_TRY(hipStreamSynchronize(st));
    // Undo the per-lane bias of 2^cbits, working modulo 3.
    const int64_t T3 = (int64_t)((1ull << cbits) % 3);
    int64_t sum3 = 0;
    for (uint64_t b = 0; b < gx; ++b) {
        const uint64_t active_lanes = std::min<uint64_t>(kMod3Block, lanes - b * kMod3Block);
        const int64_t s3 = (int64_t)((uint64_t)h_part[b] % 3);
        const int64_t bias3 = (int64_t)((active_lanes % 3) * (uint64_t)T3 % 3);
        sum3 = (sum3 + s3 - bias3 + 6) % 3;
    }
    // 2^-(n-1) = (-1)^(n-1) in F_3.
    if ((n - 1) & 1) sum3 = (3 - sum3) % 3;
    return (int)sum3;
}

}  // namespace phip
