// Fixed-sign Glynn for odd moduli q >= 2^31 (and every order the persistent
// 31-bit engine does not take): the 64-bit-word counterpart of
// glynn_generic_modp's contract (perm_c_omp.c:666-770, any odd modulus below
// 2^63, perm_c_omp.c:876-880) and of the reference's device Glynn kernels
// over 32-bit primes (perm_modp_cuda.cu:328-485).
//
// One launch per arithmetic class (wide_arith.h: Mont32, M61, PM61, Mont64) and
// register tier; blockIdx.y selects the modulus.  Each lane walks 2^cbits
// consecutive Gray states with its row sums in registers (the tier NMAX
// rows; rows at or beyond n are the identity factor, so the unrolled loops
// carry no per-row test).  The flipped column and the direction of its update
// are the same for every lane except at the middle state of the chunk,
// where the flipped bit is the chunk's lowest bit, so the deltas arrive by
// scalar loads and the select of +d or -d is scalar too.  Each wavefront
// folds its lanes' sums by cross-lane swizzles (modular adds) into one
// partial; wide_sum_kernel folds the partials per modulus.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "device_context.h"
#include "phip_arith.h"
#include "residues.h"
#include "wide.h"
#include "wide_arith.h"

namespace phip {

namespace {

constexpr int kWideBlock = 256;
constexpr int kWideLanesLog2 = 20;

template <class T>
__device__ __forceinline__ T wide_xor_lane(T v, int mask) {
    if constexpr (sizeof(T) == 4) {
        return (T)__shfl_xor((int)v, mask, 64);
    } else {
        const int lo = __shfl_xor((int)(uint32_t)v, mask, 64);
        const int hi = __shfl_xor((int)(uint32_t)((uint64_t)v >> 32), mask, 64);
        return (T)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
    }
}

// Tables hold NMAX rows per column: rs0 [q][NMAX], deltas and neg_deltas
// [q][n-1][NMAX]; rows at or beyond n are the identity factor (row sum the
// arithmetic's one, deltas 0), so the unrolled loops carry no per-row test.
// Each wavefront folds its lanes' sums and writes partials[q][wave].
template <int NMAX, class A>
__global__ __launch_bounds__(kWideBlock) void wide_glynn_kernel(const typename A::T* __restrict__ rs0,
                                                                const typename A::T* __restrict__ deltas,
                                                                const typename A::T* __restrict__ neg_deltas,
                                                                const A* __restrict__ arith, int n, int cbits,
                                                                uint64_t lanes, uint64_t waves,
                                                                uint64_t* __restrict__ partials) {
    typedef typename A::T T;
    const int q = blockIdx.y;
    const A ar = arith[q];
    const uint64_t lane = (uint64_t)blockIdx.x * kWideBlock + threadIdx.x;
    const size_t table = (size_t)q * (n - 1) * NMAX;
    const T* __restrict__ D = deltas + table;
    const T* __restrict__ ND = neg_deltas + table;

    T acc = 0;
    if (lane < lanes) {
        T rs[NMAX];
        {
            const ConstPtr<T> r0 = scalar_ptr(rs0 + (size_t)q * NMAX);
#pragma unroll
            for (int i = 0; i < NMAX; ++i) rs[i] = r0[i];
        }
        const uint64_t base = lane << cbits;
        const uint64_t gray = base ^ (base >> 1);
        // Start of the chunk: subtract the deltas of the columns set in gray(base).
        for (int j = 0; j < n - 1; ++j) {
            const bool on = (gray >> j) & 1u;
            const ConstPtr<T> nd = scalar_ptr(ND + (size_t)j * NMAX);
#pragma unroll
            for (int i = 0; i < NMAX; ++i) rs[i] = ar.add(rs[i], on ? nd[i] : T(0));
        }
        bool negative = __popcll(gray) & 1;
        const bool lane_odd = lane & 1u;

        const uint64_t states = 1ull << cbits;
        for (uint64_t t = 0; t < states; ++t) {
            if (t != 0) {
                const int j = __builtin_ctzll(t);
                if (j == cbits - 1) {
                    // Bit j of gray(base + t) is 1 xor the lane's lowest bit.
                    const bool sub = !lane_odd;
                    const ConstPtr<T> d = scalar_ptr(D + (size_t)j * NMAX);
                    const ConstPtr<T> nd = scalar_ptr(ND + (size_t)j * NMAX);
#pragma unroll
                    for (int i = 0; i < NMAX; ++i) rs[i] = ar.add(rs[i], sub ? nd[i] : d[i]);
                } else {
                    const bool sub = ((t ^ (t >> 1)) >> j) & 1u;
                    const ConstPtr<T> d = scalar_ptr((sub ? ND : D) + (size_t)j * NMAX);
#pragma unroll
                    for (int i = 0; i < NMAX; ++i) rs[i] = ar.add(rs[i], d[i]);
                }
                negative = !negative;
            }
            T prod = rs[0];
#pragma unroll
            for (int i = 1; i < NMAX; ++i) prod = ar.mul(prod, rs[i]);
            acc = ar.add(acc, negative ? ar.neg(prod) : prod);
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) acc = ar.add(acc, wide_xor_lane(acc, m));
    const uint64_t wave = lane >> 6;
    if ((threadIdx.x & 63) == 0 && wave < waves) partials[(size_t)q * waves + wave] = (uint64_t)acc;
}

// out[q] = sum of the lanes' partials modulo moduli[q].
__global__ __launch_bounds__(256) void wide_sum_kernel(const uint64_t* __restrict__ partials, uint64_t count,
                                                       const uint64_t* __restrict__ moduli,
                                                       uint64_t* __restrict__ out) {
    __shared__ uint64_t tree[256];
    const int q = blockIdx.x;
    const uint64_t p = moduli[q];
    const uint64_t* __restrict__ src = partials + (size_t)q * count;
    uint64_t s = 0;
    for (uint64_t i = threadIdx.x; i < count; i += 256) s = add_mod_u64(s, src[i], p);
    tree[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) tree[threadIdx.x] = add_mod_u64(tree[threadIdx.x], tree[threadIdx.x + w], p);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = tree[0];
}

template <class A>
A make_arith(uint64_t q);
template <>
Mont32 make_arith<Mont32>(uint64_t q) {
    return make_mont32(q);
}
template <>
M61 make_arith<M61>(uint64_t q) {
    return make_m61(q);
}
template <>
PM61 make_arith<PM61>(uint64_t q) {
    return make_pm61(q);
}
template <>
Mont64 make_arith<Mont64>(uint64_t q) {
    return make_mont64(q);
}

template <class A>
int launch_tier(int tier, dim3 grid, hipStream_t st, const typename A::T* rs0, const typename A::T* del,
                const typename A::T* ndel, const A* ar, int n, int cbits, uint64_t lanes, uint64_t waves,
                uint64_t* part) {
#define PHIP_WIDE_LAUNCH(NM)                                                                                   \
    hipLaunchKernelGGL((wide_glynn_kernel<NM, A>), grid, dim3(kWideBlock), 0, st, rs0, del, ndel, ar, n, cbits, \
                       lanes, waves, part)
    // The tables were laid out for exactly `tier` rows per column.
    if constexpr (sizeof(typename A::T) == 4) {
        switch (tier) {
            case 8: PHIP_WIDE_LAUNCH(8); break;
            case 12: PHIP_WIDE_LAUNCH(12); break;
            case 16: PHIP_WIDE_LAUNCH(16); break;
            case 20: PHIP_WIDE_LAUNCH(20); break;
            case 24: PHIP_WIDE_LAUNCH(24); break;
            case 28: PHIP_WIDE_LAUNCH(28); break;
            case 32: PHIP_WIDE_LAUNCH(32); break;
            case 36: PHIP_WIDE_LAUNCH(36); break;
            case 40: PHIP_WIDE_LAUNCH(40); break;
            case 44: PHIP_WIDE_LAUNCH(44); break;
            case 48: PHIP_WIDE_LAUNCH(48); break;
            case 56: PHIP_WIDE_LAUNCH(56); break;
            case 64: PHIP_WIDE_LAUNCH(64); break;
            default: set_error("internal: no wide kernel for this register tier"); return -1;
        }
    } else {
        switch (tier) {
            case 8: PHIP_WIDE_LAUNCH(8); break;
            case 16: PHIP_WIDE_LAUNCH(16); break;
            case 24: PHIP_WIDE_LAUNCH(24); break;
            case 32: PHIP_WIDE_LAUNCH(32); break;
            case 40: PHIP_WIDE_LAUNCH(40); break;
            case 48: PHIP_WIDE_LAUNCH(48); break;
            case 56: PHIP_WIDE_LAUNCH(56); break;
            case 64: PHIP_WIDE_LAUNCH(64); break;
            default: set_error("internal: no wide kernel for this register tier"); return -1;
        }
    }
#undef PHIP_WIDE_LAUNCH
    PHIP_HIP_TRY(hipGetLastError());
    return 0;
}

// Register tier (rows per table column) for order n: 32-bit residues every
// 4 rows from 8 to 48, then 56 and 64; 64-bit ones every 8.
int wide_tier(int n, int min_tier, bool narrow) {
    const int rows = std::max({8, n, min_tier});
    if (narrow) return rows <= 48 ? (rows + 3) / 4 * 4 : (rows + 7) / 8 * 8;
    return (rows + 7) / 8 * 8;
}

// One arithmetic class: tables, launch, fold, scale.  Buffers of the
// device context, slots 29..35 (the class runs to completion before the
// next one reuses them).
template <class A>
int run_class(DeviceContext* ctx, const int64_t* A_, int64_t stride, int n, const std::vector<WideJob>& jobs,
              std::vector<uint64_t>& out) {
    typedef typename A::T T;
    const int kk = (int)jobs.size();
    if (kk == 0) return 0;
    const int L = n - 1;
    const int W = wide_tier(n, tuning().wide_min_tier, sizeof(T) == 4);  // rows per table column
    std::vector<T> rs0((size_t)kk * W, T(0)), del((size_t)kk * (n - 1) * W, T(0)), ndel(del.size(), T(0));
    std::vector<A> ar(kk);
    std::vector<uint64_t> moduli(kk);
    for (int s = 0; s < kk; ++s) {
        const uint64_t q = jobs[s].q;
        const int64_t* M = A_ + (int64_t)jobs[s].index * stride;
        ar[s] = make_arith<A>(q);
        moduli[s] = q;
        // The arithmetic's one (R mod q in Montgomery form) for the padding rows.
        const uint64_t one = A::kRBits > 0 ? (uint64_t)((((unsigned __int128)1) << A::kRBits) % q) : 1 % q;
        for (int i = n; i < W; ++i) rs0[(size_t)s * W + i] = (T)one;
        for (int i = 0; i < n; ++i) {
            uint64_t row = 0;
            for (int j = 0; j < n; ++j) {
                const uint64_t r = (uint64_t)host_floor_mod(M[(size_t)i * n + j], (int64_t)q);
                row = add_mod_u64(row, r, q);
                if (j < n - 1) {
                    const uint64_t d = add_mod_u64(r, r, q);
                    const size_t at = ((size_t)s * (n - 1) + j) * W + i;
                    del[at] = (T)d;
                    ndel[at] = (T)(d ? q - d : 0);
                }
            }
            rs0[(size_t)s * W + i] = (T)row;
        }
    }
    const int cbits = std::min(L, std::max(std::min(L, 4), L - kWideLanesLog2));
    const uint64_t lanes = 1ull << (L - cbits);
    const uint64_t gx = (lanes + kWideBlock - 1) / kWideBlock;
    const uint64_t waves = gx * (kWideBlock / 64);

    hipStream_t st = ctx->stream();
    auto* d_rs0 = (T*)ctx->buffer(29, std::max<size_t>(rs0.size(), 1) * sizeof(T));
    auto* d_del = (T*)ctx->buffer(30, std::max<size_t>(del.size(), 1) * sizeof(T));
    auto* d_ndel = (T*)ctx->buffer(31, std::max<size_t>(ndel.size(), 1) * sizeof(T));
    auto* d_ar = (A*)ctx->buffer(32, ar.size() * sizeof(A));
    auto* d_mod = (uint64_t*)ctx->buffer(33, moduli.size() * 8);
    auto* d_part = (uint64_t*)ctx->buffer(34, (size_t)kk * waves * 8);
    auto* d_sum = (uint64_t*)ctx->buffer(35, (size_t)kk * 8);
    auto* h_sum = (uint64_t*)ctx->host_buffer(6, (size_t)kk * 8);
    if (!d_rs0 || !d_del || !d_ndel || !d_ar || !d_mod || !d_part || !d_sum || !h_sum) return -1;
    PHIP_HIP_TRY(hipMemcpyAsync(d_rs0, rs0.data(), rs0.size() * sizeof(T), hipMemcpyHostToDevice, st));
    if (!del.empty()) {
        PHIP_HIP_TRY(hipMemcpyAsync(d_del, del.data(), del.size() * sizeof(T), hipMemcpyHostToDevice, st));
        PHIP_HIP_TRY(hipMemcpyAsync(d_ndel, ndel.data(), ndel.size() * sizeof(T), hipMemcpyHostToDevice, st));
    }
    PHIP_HIP_TRY(hipMemcpyAsync(d_ar, ar.data(), ar.size() * sizeof(A), hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d_mod, moduli.data(), moduli.size() * 8, hipMemcpyHostToDevice, st));
    const bool prof = ctx->profiling;
    if (prof) PHIP_HIP_TRY(hipEventRecord(ctx->prof_start, st));
    if (launch_tier<A>(W, dim3((unsigned)gx, (unsigned)kk), st, d_rs0, d_del, d_ndel, d_ar, n, cbits, lanes, waves,
                       d_part) != 0)
        return -1;
    if (prof) PHIP_HIP_TRY(hipEventRecord(ctx->prof_stop, st));
    hipLaunchKernelGGL(wide_sum_kernel, dim3(kk), dim3(256), 0, st, d_part, waves, d_mod, d_sum);
    PHIP_HIP_TRY(hipGetLastError());
    PHIP_HIP_TRY(hipMemcpyAsync(h_sum, d_sum, (size_t)kk * 8, hipMemcpyDeviceToHost, st));
    PHIP_HIP_TRY(hipStreamSynchronize(st));
    if (prof) {
        float ms = 0;
        PHIP_HIP_TRY(hipEventElapsedTime(&ms, ctx->prof_start, ctx->prof_stop));
        ctx->profiled_ms += ms;
        ctx->profiled_launches += 1;
    }
    for (int s = 0; s < kk; ++s) {
        const uint64_t q = jobs[s].q;
        // n - 1 products each carry R^-1; Glynn's sum carries 2^(n-1).
        uint64_t scale = host_powmod((q + 1) / 2, (uint64_t)L, q);
        if (A::kRBits > 0) {
            const uint64_t r_mod_q = (uint64_t)((((unsigned __int128)1) << A::kRBits) % q);
            scale = host_mulmod(scale, host_powmod(r_mod_q, (uint64_t)L, q), q);
        }
        out[jobs[s].index] = host_mulmod(h_sum[s], scale, q);
    }
    return 0;
}

}  // namespace

int run_wide_glynn(DeviceContext* ctx, const int64_t* A, int64_t stride, int n, const std::vector<WideJob>& jobs,
                   std::vector<uint64_t>& odd_residue) {
    if (n < 2 || n > kWideMaxOrder) {
        set_error("wide Glynn orders must lie in [2, 64)");
        return -1;
    }
    std::vector<WideJob> by_class[4];
    for (const WideJob& j : jobs) {
        if (!(j.q & 1) || j.q < 3 || j.q >= (1ull << 63)) {
            set_error("wide Glynn moduli must be odd and lie in [3, 2^63)");
            return -1;
        }
        by_class[wide_class(j.q)].push_back(j);
    }
    if (run_class<Mont32>(ctx, A, stride, n, by_class[kWideMont32], odd_residue) != 0) return -1;
    if (run_class<M61>(ctx, A, stride, n, by_class[kWideM61], odd_residue) != 0) return -1;
    if (run_class<PM61>(ctx, A, stride, n, by_class[kWidePM61], odd_residue) != 0) return -1;
    if (run_class<Mont64>(ctx, A, stride, n, by_class[kWideMont64], odd_residue) != 0) return -1;
    return 0;
}

}  // namespace phip
