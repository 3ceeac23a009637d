// Coefficient-valued fixed-sign Glynn: the coefficients of per(C + x L)
// modulo odd moduli (the reference's affine_permanent_mod contract,
// perm_polynomial.py:268-330 and the native perm_polynomial_omp.c:681-790,
// 1207-1260).  Glynn's sum runs over the same 2^(n-1) Gray states as the
// integer path, with two row sums per row (the constant and the linear
// part); each state's term is the polynomial prod_i (c_i + x l_i) of degree
// n, built in registers factor by factor with one fma2 (a b + c d, one
// Montgomery factor) per coefficient, and added to the lane's n + 1
// accumulators with the state's sign.
//
// One launch per residue arithmetic (S30 below 2^30: signed lazy Montgomery
// with exact int64 accumulators; wide_arith.h: Mont31L below 2^31 with one
// lazy reduction per coefficient, Mont32, M61, PM61, Mont64) and
// register tier (every 4 rows from 8 to 40 for 32-bit residues, then 48, 56
// and 64 for moduli below 2^30, every 8 to 32 for 64-bit ones); blockIdx.y
// selects the modulus.  The tiers above 40 hold more values than the 256
// VGPRs: the compiler keeps the excess in AGPRs (no scratch; one wavefront
// per SIMD at tier 64).
// As in wide.hip the flipped column and its direction are wave-uniform
// except at a chunk's middle state, so deltas and negated deltas come by
// scalar loads.  Each wavefront folds its lanes' accumulators with
// cross-lane swizzles (modular adds), writes one partial per coefficient,
// and affine_sum_kernel folds the wavefront partials per (modulus,
// coefficient).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "affine.h"
#include "device_context.h"
#include "phip_arith.h"
#include "wide_arith.h"

namespace phip {

namespace {

constexpr int kAffineBlock = 256;
constexpr int kAffineLanesLog2 = 18;

template <class T>
__device__ __forceinline__ T xor_lane(T v, int mask) {
    if constexpr (sizeof(T) == 4) {
        return (T)__shfl_xor((int)v, mask, 64);
    } else {
        const int lo = __shfl_xor((int)(uint32_t)v, mask, 64);
        const int hi = __shfl_xor((int)(uint32_t)((uint64_t)v >> 32), mask, 64);
        return (T)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
    }
}

// Signed lazy Montgomery for odd p < 2^30 (R = 2^32).  Row sums stay in
// [0, p) as for the other arithmetics; the product coefficients are int32
// with |v| < 2^30: a b + c d < 2^61 and the reduction adds |m p| < 2^61, so
// one REDC of the exact sum returns |r| < 2^30 again with no correction
// step.  The lane accumulators are exact int64 sums (|acc| < 2^(30 + cbits)),
// reduced once at the end with one more R^-1.
struct S30 {
    typedef uint32_t T;
    static constexpr int kRBits = 32;
    uint32_t p;
    int32_t negp;
    uint32_t pinv;  // p^-1 mod 2^32
    __device__ __forceinline__ T add(T a, T b) const {
        const T s = a + b;  // < 2^31
        return s >= p ? s - p : s;
    }
    __device__ __forceinline__ T neg(T a) const { return a ? p - a : 0; }
    __device__ __forceinline__ int32_t redc(int64_t t) const {
        const int32_t m = (int32_t)((uint32_t)t * pinv);
        return (int32_t)((t + (int64_t)m * (int64_t)negp) >> 32);
    }
};

// How a term is built and summed for arithmetic A: by default the product
// coefficients and the accumulators live in A's residues.  kCompact marks the
// register tiers above 32 rows, where the accumulators must take one
// register per coefficient.
template <class A, bool kCompact = false>
struct AffineOps {
    typedef typename A::T T;
    typedef typename A::T V;
    typedef typename A::T Acc;
    static constexpr int kExtraR = 0;
    __device__ __forceinline__ static V from_row(T x) { return x; }
    __device__ __forceinline__ static V mul(const A& ar, V a, T b) { return ar.mul(a, b); }
    __device__ __forceinline__ static V fma2(const A& ar, V a, T b, V c, T d) { return ar.fma2(a, b, c, d); }
    __device__ __forceinline__ static Acc acc_add(const A& ar, Acc acc, V v, bool negative) {
        return ar.add(acc, negative ? ar.neg(v) : v);
    }
    __device__ __forceinline__ static T finish(const A&, Acc acc) { return acc; }
};

template <>
struct AffineOps<S30, false> {
    typedef uint32_t T;
    typedef int32_t V;
    typedef int64_t Acc;
    static constexpr int kExtraR = 1;
    __device__ __forceinline__ static V from_row(T x) { return (int32_t)x; }
    __device__ __forceinline__ static V mul(const S30& ar, V a, T b) { return ar.redc((int64_t)a * (int32_t)b); }
    __device__ __forceinline__ static V fma2(const S30& ar, V a, T b, V c, T d) {
        return ar.redc((int64_t)a * (int32_t)b + (int64_t)c * (int32_t)d);
    }
    __device__ __forceinline__ static Acc acc_add(const S30&, Acc acc, V v, bool negative) {
        return negative ? acc - v : acc + v;
    }
    __device__ __forceinline__ static T finish(const S30& ar, Acc acc) {
        const int32_t r = ar.redc(acc) % (int32_t)ar.p;
        return (T)(r < 0 ? r + (int32_t)ar.p : r);
    }
};

// S30 above 32 rows: the same signed lazy products, but each accumulator is
// one int32 kept in (-p, p) (|acc +- v| < 2p < 2^31, then one correction),
// so 41 coefficients fit beside 80 row sums.  The accumulators hold plain
// residue sums of the products, so no extra R^-1 is removed at the end.
template <>
struct AffineOps<S30, true> {
    typedef uint32_t T;
    typedef int32_t V;
    typedef int32_t Acc;
    static constexpr int kExtraR = 0;
    __device__ __forceinline__ static V from_row(T x) { return (int32_t)x; }
    __device__ __forceinline__ static V mul(const S30& ar, V a, T b) { return ar.redc((int64_t)a * (int32_t)b); }
    __device__ __forceinline__ static V fma2(const S30& ar, V a, T b, V c, T d) {
        return ar.redc((int64_t)a * (int32_t)b + (int64_t)c * (int32_t)d);
    }
    __device__ __forceinline__ static Acc acc_add(const S30& ar, Acc acc, V v, bool negative) {
        const int32_t p = (int32_t)ar.p;
        int32_t s = negative ? acc - v : acc + v;
        s = s >= p ? s - p : s;
        return s <= -p ? s + p : s;
    }
    __device__ __forceinline__ static T finish(const S30& ar, Acc acc) {
        return (T)(acc < 0 ? acc + (int32_t)ar.p : acc);
    }
};

// Tables hold NMAX rows per column: c0, l0: [q][NMAX] state-0 row sums;
// dc, ndc, dl, ndl: [q][n-1][NMAX] deltas (2 C, 2 L by column) and their
// negations.  Rows at or beyond n are the identity factor (constant part
// the arithmetic's one, linear part 0, deltas 0), so the unrolled loops run
// without per-row tests and the product is unchanged by them.  partials:
// [q][waves][n + 1].
template <int NMAX, class A>
__global__ __launch_bounds__(kAffineBlock) void affine_glynn_kernel(
    const typename A::T* __restrict__ c0, const typename A::T* __restrict__ l0,
    const typename A::T* __restrict__ dc, const typename A::T* __restrict__ ndc,
    const typename A::T* __restrict__ dl, const typename A::T* __restrict__ ndl, const A* __restrict__ arith, int n,
    int cbits, uint64_t lanes, uint64_t waves, uint64_t* __restrict__ partials) {
    typedef typename A::T T;
    typedef AffineOps<A, (NMAX > 32)> Ops;
    typedef typename Ops::V V;
    const int q = blockIdx.y;
    const A ar = arith[q];
    const uint64_t lane = (uint64_t)blockIdx.x * kAffineBlock + threadIdx.x;
    const bool active = lane < lanes;
    const size_t table = (size_t)q * (n - 1) * NMAX;

    typename Ops::Acc acc[NMAX + 1];
#pragma unroll
    for (int d = 0; d <= NMAX; ++d) acc[d] = 0;

    if (active) {
        T c[NMAX], l[NMAX];
        {
            const ConstPtr<T> pc = scalar_ptr(c0 + (size_t)q * NMAX);
            const ConstPtr<T> pl = scalar_ptr(l0 + (size_t)q * NMAX);
#pragma unroll
            for (int i = 0; i < NMAX; ++i) {
                c[i] = pc[i];
                l[i] = pl[i];
            }
        }
        const uint64_t base = lane << cbits;
        const uint64_t gray = base ^ (base >> 1);
        for (int j = 0; j < n - 1; ++j) {
            const bool on = (gray >> j) & 1u;
            const ConstPtr<T> nc = scalar_ptr(ndc + table + (size_t)j * NMAX);
            const ConstPtr<T> nl = scalar_ptr(ndl + table + (size_t)j * NMAX);
#pragma unroll
            for (int i = 0; i < NMAX; ++i) {
                c[i] = ar.add(c[i], on ? nc[i] : T(0));
                l[i] = ar.add(l[i], on ? nl[i] : T(0));
            }
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
                    const ConstPtr<T> pc = scalar_ptr(dc + table + (size_t)j * NMAX);
                    const ConstPtr<T> nc = scalar_ptr(ndc + table + (size_t)j * NMAX);
                    const ConstPtr<T> pl = scalar_ptr(dl + table + (size_t)j * NMAX);
                    const ConstPtr<T> nl = scalar_ptr(ndl + table + (size_t)j * NMAX);
#pragma unroll
                    for (int i = 0; i < NMAX; ++i) {
                        c[i] = ar.add(c[i], sub ? nc[i] : pc[i]);
                        l[i] = ar.add(l[i], sub ? nl[i] : pl[i]);
                    }
                } else {
                    const bool sub = ((t ^ (t >> 1)) >> j) & 1u;
                    const ConstPtr<T> pc = scalar_ptr((sub ? ndc : dc) + table + (size_t)j * NMAX);
                    const ConstPtr<T> pl = scalar_ptr((sub ? ndl : dl) + table + (size_t)j * NMAX);
#pragma unroll
                    for (int i = 0; i < NMAX; ++i) {
                        c[i] = ar.add(c[i], pc[i]);
                        l[i] = ar.add(l[i], pl[i]);
                    }
                }
                negative = !negative;
            }
            // prod_i (c_i + x l_i), one factor at a time, highest degree first.
            V P[NMAX + 1];
            P[0] = Ops::from_row(c[0]);
            P[1] = Ops::from_row(l[0]);
#pragma unroll
            for (int i = 1; i < NMAX; ++i) {
                P[i + 1] = Ops::mul(ar, P[i], l[i]);
#pragma unroll
                for (int d = i; d >= 1; --d) P[d] = Ops::fma2(ar, P[d], c[i], P[d - 1], l[i]);
                P[0] = Ops::mul(ar, P[0], c[i]);
            }
#pragma unroll
            for (int d = 0; d <= NMAX; ++d) acc[d] = Ops::acc_add(ar, acc[d], P[d], negative);
        }
    }

    // Wavefront fold of every coefficient, then one partial per wavefront.
    const uint64_t wave = lane >> 6;
#pragma unroll
    for (int d = 0; d <= NMAX; ++d) {
        if (d <= n) {
            T v = Ops::finish(ar, acc[d]);
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) v = ar.add(v, xor_lane(v, m));
            if ((threadIdx.x & 63) == 0 && wave < waves) partials[((size_t)q * waves + wave) * (n + 1) + d] = (uint64_t)v;
        }
    }
}

// out[q][d] = sum over wavefronts of partials[q][w][d] modulo moduli[q].
__global__ __launch_bounds__(256) void affine_sum_kernel(const uint64_t* __restrict__ partials, uint64_t waves,
                                                         int width, const uint64_t* __restrict__ moduli,
                                                         uint64_t* __restrict__ out) {
    __shared__ uint64_t tree[256];
    const int q = blockIdx.y;
    const int d = blockIdx.x;
    const uint64_t p = moduli[q];
    uint64_t s = 0;
    for (uint64_t w = threadIdx.x; w < waves; w += 256)
        s = add_mod_u64(s, partials[((size_t)q * waves + w) * width + d], p);
    tree[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) tree[threadIdx.x] = add_mod_u64(tree[threadIdx.x], tree[threadIdx.x + k], p);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)q * width + d] = tree[0];
}

template <class A>
A make_affine_arith(uint64_t q);
template <>
S30 make_affine_arith<S30>(uint64_t q) {
    return S30{(uint32_t)q, -(int32_t)q, inverse_mod_2_32((uint32_t)q)};
}
template <>
Mont31L make_affine_arith<Mont31L>(uint64_t q) {
    return make_mont31l(q);
}
template <>
Mont32 make_affine_arith<Mont32>(uint64_t q) {
    return make_mont32(q);
}
template <>
M61 make_affine_arith<M61>(uint64_t q) {
    return make_m61(q);
}
template <>
PM61 make_affine_arith<PM61>(uint64_t q) {
    return make_pm61(q);
}
template <>
Mont64 make_affine_arith<Mont64>(uint64_t q) {
    return make_mont64(q);
}

template <class A>
int launch_affine(int tier, dim3 grid, hipStream_t st, const typename A::T* c0, const typename A::T* l0,
                  const typename A::T* dc, const typename A::T* ndc, const typename A::T* dl,
                  const typename A::T* ndl, const A* ar, int n, int cbits, uint64_t lanes, uint64_t waves,
                  uint64_t* part) {
#define PHIP_AFFINE_LAUNCH(NM)                                                                                  \
    hipLaunchKernelGGL((affine_glynn_kernel<NM, A>), grid, dim3(kAffineBlock), 0, st, c0, l0, dc, ndc, dl, ndl, \
                       ar, n, cbits, lanes, waves, part)
    // The tables were laid out for exactly `tier` rows per column.
    if constexpr (sizeof(typename A::T) == 4) {
        switch (tier) {
            case 8: PHIP_AFFINE_LAUNCH(8); break;
            case 12: PHIP_AFFINE_LAUNCH(12); break;
            case 16: PHIP_AFFINE_LAUNCH(16); break;
            case 20: PHIP_AFFINE_LAUNCH(20); break;
            case 24: PHIP_AFFINE_LAUNCH(24); break;
            case 28: PHIP_AFFINE_LAUNCH(28); break;
            case 32: PHIP_AFFINE_LAUNCH(32); break;
            case 36: PHIP_AFFINE_LAUNCH(36); break;
            case 40: PHIP_AFFINE_LAUNCH(40); break;
            case 48:
            case 56:
            case 64:
                if constexpr (std::is_same<A, S30>::value) {
                    if (tier == 48) PHIP_AFFINE_LAUNCH(48);
                    else if (tier == 56) PHIP_AFFINE_LAUNCH(56);
                    else PHIP_AFFINE_LAUNCH(64);
                    break;
                } else {
                    set_error("internal: no affine kernel for this register tier");
                    return -1;
                }
            default: set_error("internal: no affine kernel for this register tier"); return -1;
        }
    } else {
        switch (tier) {
            case 8: PHIP_AFFINE_LAUNCH(8); break;
            case 16: PHIP_AFFINE_LAUNCH(16); break;
            case 24: PHIP_AFFINE_LAUNCH(24); break;
            case 32: PHIP_AFFINE_LAUNCH(32); break;
            default: set_error("internal: no affine kernel for this register tier"); return -1;
        }
    }
#undef PHIP_AFFINE_LAUNCH
    PHIP_HIP_TRY(hipGetLastError());
    return 0;
}

// One arithmetic class.  Device buffers 36..39 of the context plus
// host buffer 7; the class completes before the next reuses them.
template <class A>
int run_affine_class(DeviceContext* ctx, const int64_t* C, const int64_t* L, int64_t stride, int n,
                     const std::vector<AffineJob>& jobs, int min_tier, int64_t* out) {
    typedef typename A::T T;
    const int kk = (int)jobs.size();
    if (kk == 0) return 0;
    // Register tier: the smallest that holds n rows (32-bit residues have
    // tiers every 4 rows from 8 to 40, and moduli below 2^30 every 8 from
    // 48 to 64; 64-bit ones every 8 to 32), or the tuning floor.
    constexpr bool kS30 = std::is_same<A, S30>::value;
    const int top = sizeof(T) == 4 ? (kS30 ? kAffineMaxOrderS30 + 1 : kAffineMaxOrder) : kAffineMaxOrderWide;
    if (n > top) {
        set_error(sizeof(T) == 4 ? "affine permanents of order above 40 need moduli below 2^30 on the device"
                                 : "affine permanents of order above 32 need moduli below 2^32 on the device");
        return -1;
    }
    const int rows = std::min(top, std::max({8, n, min_tier}));
    const int step = (sizeof(T) == 4 && rows <= kAffineMaxOrder) ? 4 : 8;
    const int W = std::min(top, (rows + step - 1) / step * step);  // rows per table column
    const size_t cols = (size_t)(n - 1) * W;
    // One host image: c0, l0 ([kk][W] each), then dc, ndc, dl, ndl
    // ([kk][n-1][W] each); rows n..W-1 are the identity factor.
    std::vector<T> img((size_t)kk * (2 * W + 4 * cols), T(0));
    T* h_c0 = img.data();
    T* h_l0 = h_c0 + (size_t)kk * W;
    T* h_dc = h_l0 + (size_t)kk * W;
    T* h_ndc = h_dc + (size_t)kk * cols;
    T* h_dl = h_ndc + (size_t)kk * cols;
    T* h_ndl = h_dl + (size_t)kk * cols;
    std::vector<A> ar(kk);
    std::vector<uint64_t> moduli(kk);
    for (int s = 0; s < kk; ++s) {
        const uint64_t q = jobs[s].q;
        const int64_t* Cs = C + (int64_t)jobs[s].index * stride;
        const int64_t* Ls = L + (int64_t)jobs[s].index * stride;
        ar[s] = make_affine_arith<A>(q);
        moduli[s] = q;
        // The arithmetic's one: R mod q in Montgomery form, else 1.
        const uint64_t one = A::kRBits > 0 ? (uint64_t)((((unsigned __int128)1) << A::kRBits) % q) : 1 % q;
        for (int i = 0; i < W; ++i) {
            if (i >= n) {
                h_c0[(size_t)s * W + i] = (T)one;
                continue;
            }
            uint64_t rc = 0, rl = 0;
            for (int j = 0; j < n; ++j) {
                const uint64_t vc = (uint64_t)host_floor_mod(Cs[(size_t)i * n + j], (int64_t)q);
                const uint64_t vl = (uint64_t)host_floor_mod(Ls[(size_t)i * n + j], (int64_t)q);
                rc = add_mod_u64(rc, vc, q);
                rl = add_mod_u64(rl, vl, q);
                if (j < n - 1) {
                    const size_t at = (size_t)s * cols + (size_t)j * W + i;
                    const uint64_t d2c = add_mod_u64(vc, vc, q), d2l = add_mod_u64(vl, vl, q);
                    h_dc[at] = (T)d2c;
                    h_ndc[at] = (T)(d2c ? q - d2c : 0);
                    h_dl[at] = (T)d2l;
                    h_ndl[at] = (T)(d2l ? q - d2l : 0);
                }
            }
            h_c0[(size_t)s * W + i] = (T)rc;
            h_l0[(size_t)s * W + i] = (T)rl;
        }
    }
    const int Lbits = n - 1;
    const int lanes_log2 = std::min(Lbits, kAffineLanesLog2);
    const int cbits = Lbits - lanes_log2;
    const uint64_t lanes = 1ull << lanes_log2;
    const uint64_t gx = (lanes + kAffineBlock - 1) / kAffineBlock;
    const uint64_t waves = gx * (kAffineBlock / 64);
    const int tier = W;
    const int width = n + 1;

    hipStream_t st = ctx->stream();
    auto* d_img = (T*)ctx->buffer(36, img.size() * sizeof(T));
    auto* d_ar = (A*)ctx->buffer(37, ar.size() * sizeof(A));
    auto* d_mod = (uint64_t*)ctx->buffer(38, moduli.size() * 8);
    auto* d_part = (uint64_t*)ctx->buffer(39, (size_t)kk * waves * width * 8 + (size_t)kk * width * 8);
    auto* h_out = (uint64_t*)ctx->host_buffer(7, (size_t)kk * width * 8);
    if (!d_img || !d_ar || !d_mod || !d_part || !h_out) return -1;
    uint64_t* d_sum = d_part + (size_t)kk * waves * width;
    PHIP_HIP_TRY(hipMemcpyAsync(d_img, img.data(), img.size() * sizeof(T), hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d_ar, ar.data(), ar.size() * sizeof(A), hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d_mod, moduli.data(), moduli.size() * 8, hipMemcpyHostToDevice, st));
    const T* d_c0 = d_img;
    const T* d_l0 = d_c0 + (size_t)kk * W;
    const T* d_dc = d_l0 + (size_t)kk * W;
    const T* d_ndc = d_dc + (size_t)kk * cols;
    const T* d_dl = d_ndc + (size_t)kk * cols;
    const T* d_ndl = d_dl + (size_t)kk * cols;
    const bool prof = ctx->profiling;
    if (prof) PHIP_HIP_TRY(hipEventRecord(ctx->prof_start, st));
    if (launch_affine<A>(tier, dim3((unsigned)gx, (unsigned)kk), st, d_c0, d_l0, d_dc, d_ndc, d_dl, d_ndl, d_ar, n,
                         cbits, lanes, waves, d_part) != 0)
        return -1;
    if (prof) PHIP_HIP_TRY(hipEventRecord(ctx->prof_stop, st));
    hipLaunchKernelGGL(affine_sum_kernel, dim3((unsigned)width, (unsigned)kk), dim3(256), 0, st, d_part, waves,
                       width, d_mod, d_sum);
    PHIP_HIP_TRY(hipGetLastError());
    PHIP_HIP_TRY(hipMemcpyAsync(h_out, d_sum, (size_t)kk * width * 8, hipMemcpyDeviceToHost, st));
    PHIP_HIP_TRY(hipStreamSynchronize(st));
    if (prof) {
        float ms = 0;
        PHIP_HIP_TRY(hipEventElapsedTime(&ms, ctx->prof_start, ctx->prof_stop));
        ctx->profiled_ms += ms;
        ctx->profiled_launches += 1;
    }
    for (int s = 0; s < kk; ++s) {
        const uint64_t q = jobs[s].q;
        // n - 1 Montgomery factors per term (one more for the lane
        // reduction of exact accumulators); Glynn's sum carries 2^(n-1).
        uint64_t scale = host_powmod((q + 1) / 2, (uint64_t)Lbits, q);
        if (A::kRBits > 0) {
            const uint64_t r_mod_q = (uint64_t)((((unsigned __int128)1) << A::kRBits) % q);
            const int extra = W > 32 ? AffineOps<A, true>::kExtraR : AffineOps<A, false>::kExtraR;
            scale = host_mulmod(scale, host_powmod(r_mod_q, (uint64_t)(Lbits + extra), q), q);
        }
        for (int d = 0; d < width; ++d)
            out[(size_t)jobs[s].index * width + d] = (int64_t)host_mulmod(h_out[(size_t)s * width + d], scale, q);
    }
    return 0;
}

}  // namespace

int affine_permanent_residues(DeviceContext* ctx, const int64_t* C, const int64_t* L, int64_t stride, int n,
                              const int64_t* moduli, int k, int min_tier, int64_t* out) {
    if (n < 0 || n > kAffineMaxOrderS30) {
        set_error("affine permanent orders must lie in [0, 63] on the device");
        return -1;
    }
    if (k < 0 || (k > 0 && (!moduli || !out))) {
        set_error("invalid modulus list");
        return -1;
    }
    if (n > 0 && (!C || !L)) {
        set_error("matrix pointer is null");
        return -1;
    }
    for (int i = 0; i < k; ++i) {
        const int64_t m = moduli[i];
        if (m < 3 || !(m & 1)) {
            set_error("affine moduli must be odd and at least 3");
            return -1;
        }
    }
    const int width = n + 1;
    if (n == 0) {
        for (int i = 0; i < k; ++i) out[i] = 1;
        return 0;
    }
    if (n == 1) {
        for (int i = 0; i < k; ++i) {
            out[(size_t)i * 2] = host_floor_mod(C[(int64_t)i * stride], moduli[i]);
            out[(size_t)i * 2 + 1] = host_floor_mod(L[(int64_t)i * stride], moduli[i]);
        }
        return 0;
    }
    std::vector<AffineJob> by_class[6];
    for (int i = 0; i < k; ++i) {
        const uint64_t q = (uint64_t)moduli[i];
        int cls;
        if (q < (1ull << 30)) {
            cls = 5;
        } else if (q < (1ull << 31)) {
            cls = 0;
        } else {
            switch (wide_class(q)) {
                case kWideMont32: cls = 1; break;
                case kWideM61: cls = 2; break;
                case kWidePM61: cls = 3; break;
                default: cls = 4; break;
            }
        }
        by_class[cls].push_back({i, q});
    }
    (void)width;
    if (run_affine_class<S30>(ctx, C, L, stride, n, by_class[5], min_tier, out) != 0) return -1;
    if (run_affine_class<Mont31L>(ctx, C, L, stride, n, by_class[0], min_tier, out) != 0) return -1;
    if (run_affine_class<Mont32>(ctx, C, L, stride, n, by_class[1], min_tier, out) != 0) return -1;
    if (run_affine_class<M61>(ctx, C, L, stride, n, by_class[2], min_tier, out) != 0) return -1;
    if (run_affine_class<PM61>(ctx, C, L, stride, n, by_class[3], min_tier, out) != 0) return -1;
    if (run_affine_class<Mont64>(ctx, C, L, stride, n, by_class[4], min_tier, out) != 0) return -1;
    return 0;
}

}  // namespace phip
