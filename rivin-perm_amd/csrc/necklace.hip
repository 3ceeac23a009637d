// Schur cofactor minors by necklace orbits (gfx950).
//
// For a prime order n and omega of order dividing n, reindexing the (0,0)
// minor of S_n by powers of a primitive root g mod n makes it circulant:
// minor'[r][c] = w[(c - r) mod L], w[k] = omega^(g^k mod n), L = n - 1 (the
// reference's construction, perm_c_omp.c:193-215).  Glynn's sum over all
// 2^L column-sign words of a circulant matrix is invariant under rotating
// the word and under complementing it (L is even), so it is the sum over
// orbit representatives weighted by orbit size (perm_c_omp.c:79-136,
// 249-284).  Representatives are the necklaces (least rotations) of weight
// at most L/2: weight < L/2 stands for its rotations and their complements
// (2 * period words), weight == L/2 for its rotations only (period words; a
// complement that is not a rotation is another weight-L/2 necklace and is
// counted on its own).  This covers every word exactly once.
//
// Enumeration.  Necklaces come from the Fredricksen-Kessler-Maiorana
// successor on prenecklaces, kept as L-bit words (a[1] in the top bit):
// set the last 0 bit at position i, then repeat the first i bits; the
// result is a necklace when i divides L.  A word whose weight exceeds L/2
// has its tail after the (L/2+1)-th set bit filled with ones first, so the
// successor skips every over-weight word with that prefix (the pruning of
// the reference's recursion, perm_c_omp.c:138-157).  The host cuts the FKM
// tree at depth D >= L/2 into prefix tasks; each lane takes tasks from a
// counter and walks its subtree, so every lane emits one representative
// per step of the wavefront's loop.
//
// Row sums.  Each lane keeps the exact integer row sums of its last emitted
// word (centred entries, |sum| < L p / 2 < 2^37) and updates them by the
// bits that differ from the next one.  The lanes' changed columns are
// merged into one wave-uniform set, so each column's entries w[(c - r) mod
// L] come from scalar loads and each row update is one v_mad_i64_i32 with a
// per-lane factor in {-2, 0, +2}.  The product reduces each row once
// (Montgomery, R = 2^32) and multiplies the rows in a balanced tree, so a
// term carries R^-(2L-1); folding a lane's accumulator adds one more R^-1.
// Two arithmetics share the kernel: signed Montgomery for odd p < 2^31 and
// unsigned Montgomery for odd p < 2^32 (the reference's 32-bit primes
// q = 1 mod n, perm_cuda.py:289-302, perm_modp_cuda.cu:487-701).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "device_context.h"
#include "glynn31.h"
#include "necklace.h"
#include "phip_arith.h"
#include "phip_reduce.h"
#include "residues.h"
#include "wide_arith.h"

namespace phip {

namespace {

constexpr int kNecklaceBlock = 256;
// Representatives per lane between accumulator folds: |term * orbit| <
// 2^31 * 92 < 2^38, so 2^22 of them stay below 2^60.
constexpr uint32_t kFoldReps = 1u << 22;
// Eighths of the lanes with work left that must hold a representative
// before a wavefront evaluates its terms.
constexpr int kReadyEighths = 6;

// The first L bits of prefix, prefix, prefix, ... (prefix has i bits).
template <int L>
__device__ __forceinline__ uint64_t periodic_extend(uint64_t prefix, int i) {
    uint64_t y = prefix;
    int l = i;
    while (l < L) {
        if (2 * l <= L) {
            y = (y << l) | y;
            l <<= 1;
        } else {
            y = (y << (L - l)) | (y >> (2 * l - L));
            l = L;
        }
    }
    return y;
}

// Bit d set for every divisor d of m (m < 64): a prenecklace of period d is
// a necklace exactly when d divides the length.
constexpr uint64_t divisor_mask(int m) {
    uint64_t r = 0;
    for (int d = 1; d <= m; ++d)
        if (m % d == 0) r |= 1ull << d;
    return r;
}

__device__ __forceinline__ uint64_t wave_or_u64(uint64_t v) {
    v |= xor_lanes_u64(v, 1);
    v |= xor_lanes_u64(v, 2);
    v |= xor_lanes_u64(v, 4);
    v |= xor_lanes_u64(v, 8);
    v |= xor_lanes_u64(v, 16);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 0) |
                        (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 32);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 0) |
                        (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 32);
    return ((uint64_t)hi << 32) | lo;
}

// Signed Montgomery, odd p < 2^31: values in (-2^31, 2^31).
struct NeckSigned31 {
    typedef int32_t T;
    int32_t p;
    int32_t negp;
    uint32_t pinv;
    __device__ __forceinline__ T row(int64_t s) const { return redc31(s, negp, pinv); }
    __device__ __forceinline__ T mul(T a, T b) const { return mont31(a, b, negp, pinv); }
    __device__ __forceinline__ int32_t centred(T v) const { return v; }
    __device__ __forceinline__ int64_t fold(int64_t acc) const { return redc31(acc, negp, pinv); }
};

// Unsigned Montgomery, odd p < 2^32: values in [0, p).  Row sums (|s| <
// 23 p) and accumulators (|acc| < 2^60) are made non-negative by adding a
// multiple of p that keeps them below p 2^32, then reduced by REDC.
struct NeckUnsigned32 {
    typedef uint32_t T;
    Mont32 m;
    int64_t row_bias;  // 32 p
    int64_t acc_bias;  // 2^30 p >= 2^61
    __device__ __forceinline__ T redc(uint64_t x) const {
        const uint32_t k = (uint32_t)x * m.np;
        const uint64_t s = x + (uint64_t)k * m.p;
        const uint64_t r = (s >> 32) | ((uint64_t)(s < x) << 32);
        return (T)(r >= m.p ? r - m.p : r);
    }
    __device__ __forceinline__ T row(int64_t s) const { return redc((uint64_t)(s + row_bias)); }
    __device__ __forceinline__ T mul(T a, T b) const { return m.mul(a, b); }
    __device__ __forceinline__ int32_t centred(T v) const { return (int32_t)(v > (m.p >> 1) ? v - m.p : v); }
    __device__ __forceinline__ int64_t fold(int64_t acc) const { return (int64_t)redc((uint64_t)(acc + acc_bias)); }
};

// Stand-in arithmetic of the counting walk: the walk returns before any
// term is formed, so these are never executed.
struct NeckCount {
    typedef int32_t T;
    __device__ __forceinline__ T row(int64_t) const { return 0; }
    __device__ __forceinline__ T mul(T, T) const { return 0; }
    __device__ __forceinline__ int32_t centred(T v) const { return v; }
    __device__ __forceinline__ int64_t fold(int64_t) const { return 0; }
};

// tasks[t] = prefix bits (depth bits, a[1] on top) | period << 56.
// E[j] = w[(-j) mod L] centred, 0 <= j < 2L, so row r of column c reads
// E[L - c + r].  block_partials[b] receives the block's sum of lane
// residues in [0, p).  With kCount the walk forms no terms: block_partials[b]
// receives the number of orbits under rotation and complement the block's
// representatives make up and block_partials[gridDim.x + b] the number of
// words they stand for.
template <int L, class Ar, bool kCount>
__global__ __launch_bounds__(kNecklaceBlock) void necklace_kernel(const uint64_t* __restrict__ tasks,
                                                                  uint32_t n_tasks, int depth,
                                                                  const int32_t* __restrict__ E, int64_t base_sum,
                                                                  Ar ar, int64_t p,
                                                                  uint32_t* __restrict__ counter,
                                                                  uint64_t* __restrict__ block_partials) {
    typedef typename Ar::T T;
    constexpr uint64_t kMask = (1ull << L) - 1;
    constexpr uint64_t kDivisors = divisor_mask(L);
    int64_t rs[L];
#pragma unroll
    for (int r = 0; r < L; ++r) rs[r] = base_sum;
    uint64_t prev = 0, w = 0;
    int period = 1;
    bool valid = false, done = false, have = false;
    int64_t acc = 0, total = 0;
    uint32_t reps = 0;

    for (;;) {
        // Advance.  Each iteration, every lane that holds no representative
        // and has work left takes one step of its walk: a new task, or the
        // successor of its last word.  The wavefront moves on to the terms
        // once 6/8 of the lanes with work left hold a representative, so a
        // lane in a long run of skipped words (periods that do not divide L,
        // over-weight prefixes, task changes) does not hold the others back;
        // it keeps its place and goes on in the next round.  Every
        // iteration moves each waiting lane one step, so the loop ends.
        for (;;) {
            if (!have && !done) {
                bool candidate = false;
                if (!valid) {
                    const uint32_t t = atomicAdd(counter, 1u);
                    if (t >= n_tasks) {
                        done = true;
                    } else {
                        const uint64_t task = tasks[t];
                        period = (int)(task >> 56);
                        const uint64_t bits = task & ((1ull << 56) - 1);
                        w = periodic_extend<L>(bits >> (depth - period), period);
                        valid = true;
                        candidate = true;
                    }
                } else {
                    uint64_t x = w;
                    const int wt = __popcll(x);
                    if (2 * wt > L) {
                        // Every word sharing the prefix up to the (L/2+1)-th
                        // set bit is over weight: fill the rest with ones so
                        // the successor skips that whole subtree.
                        uint64_t y = x;
                        for (int k = L / 2 + 1; k < wt; ++k) y &= y - 1;
                        x |= (1ull << __builtin_ctzll(y)) - 1;
                    }
                    const uint64_t z = ~x & kMask;
                    const int b = z ? __builtin_ctzll(z) : 0;
                    const int i = L - b;
                    if (z == 0 || i <= depth) {
                        // The walk leaves this task's subtree.
                        valid = false;
                    } else {
                        const uint64_t pre = (x >> b) | 1u;
                        // depth >= L/2, so i > L/2 and one shift-or extends the prefix.
                        w = (2 * i >= L) ? ((pre << (L - i)) | (pre >> (2 * i - L))) : periodic_extend<L>(pre, i);
                        period = i;
                        candidate = true;
                    }
                }
                if (candidate && ((kDivisors >> period) & 1u) && 2 * __popcll(w) <= L) have = true;
            }
            // Lanes with work left, and those among them holding a term.
            const int active = __popcll(__ballot(!done));
            const int holding = __popcll(__ballot(have));
            if (kReadyEighths * active <= 8 * holding) break;
        }
        if (__ballot(!done) == 0) break;

        if constexpr (kCount) {
            if (have) {
                const int wt = __popcll(w);
                if (2 * wt < L) {
                    acc += 1;
                    total += 2 * period;
                } else {
                    // Weight L/2: the walk visits a necklace and the necklace
                    // of its complement separately, each for its rotations.
                    // They form one orbit under rotation and complement,
                    // counted once, at the smaller of the two (numeric order
                    // of the L-bit words is lexicographic order); a necklace
                    // whose complement is one of its rotations is its own
                    // pair.
                    const uint64_t c = ~w & kMask;
                    uint64_t least = c;
                    for (int r = 1; r < L; ++r) {
                        const uint64_t rot = ((c << r) | (c >> (L - r))) & kMask;
                        least = rot < least ? rot : least;
                    }
                    acc += (w <= least) ? 1 : 0;
                    total += period;
                }
            }
            have = false;
            continue;
        }

        // Row sums of the new representative: one uniform pass per column
        // that changed in any lane.
        const uint64_t diff = have ? (prev ^ w) : 0;
        uint64_t u = wave_or_u64(diff);
        while (u) {
            const int b = __builtin_ctzll(u);
            u &= u - 1;
            const int col = L - 1 - b;
            const int32_t sgn = ((diff >> b) & 1u) ? (((w >> b) & 1u) ? -2 : 2) : 0;
            const ConstCol e = opaque_ptr(E + (L - col));
#pragma unroll
            for (int r = 0; r < L; ++r) rs[r] += (int64_t)e[r] * (int64_t)sgn;
        }
        if (have) prev = w;

        // Term: reduce each row, multiply in a balanced tree.
        T v[L];
#pragma unroll
        for (int r = 0; r < L; ++r) v[r] = ar.row(rs[r]);
#pragma unroll
        for (int s = 1; s < L; s *= 2) {
#pragma unroll
            for (int r = 0; r + s < L; r += 2 * s) v[r] = ar.mul(v[r], v[r + s]);
        }
        const int wt = __popcll(w);
        const int orbit = (2 * wt < L) ? 2 * period : period;
        const int32_t weight = have ? ((wt & 1) ? -orbit : orbit) : 0;
        acc += (int64_t)ar.centred(v[0]) * (int64_t)weight;
        if (++reps == kFoldReps) {
            total += ar.fold(acc);
            acc = 0;
            reps = 0;
        }
        have = false;
    }
    if constexpr (kCount) {
        const uint64_t reps_sum = block_reduce_sum_u64<kNecklaceBlock>((uint64_t)acc);
        const uint64_t words_sum = block_reduce_sum_u64<kNecklaceBlock>((uint64_t)total);
        if (threadIdx.x == 0) {
            block_partials[blockIdx.x] = reps_sum;
            block_partials[gridDim.x + blockIdx.x] = words_sum;
        }
    } else {
        total += ar.fold(acc);
        int64_t r = total % p;
        if (r < 0) r += p;
        const uint64_t s = block_reduce_sum_u64<kNecklaceBlock>((uint64_t)r);
        if (threadIdx.x == 0) block_partials[blockIdx.x] = s;
    }
}

// FKM prefix tasks of the given depth in lexicographic order (subtrees with
// long runs of zeros, the largest, come first).
void collect_tasks(int L, int depth, std::vector<uint64_t>& out) {
    struct Frame {
        int t, period;
        uint64_t prefix;
        int stage;
    };
    std::vector<Frame> stack;
    stack.push_back({1, 1, 0, 0});
    while (!stack.empty()) {
        Frame& f = stack.back();
        if (f.t > depth) {
            out.push_back(f.prefix | ((uint64_t)f.period << 56));
            stack.pop_back();
            continue;
        }
        // a[t - period] is bit (t-1) - (t-period) = period - 1 of the (t-1)-bit prefix, counted from the low end.
        const int bit = (int)((f.prefix >> (f.period - 1)) & 1u);
        if (f.stage == 0) {
            f.stage = 1;
            const Frame next{f.t + 1, f.period, (f.prefix << 1) | (uint64_t)bit, 0};
            stack.push_back(next);
        } else if (f.stage == 1) {
            f.stage = 2;
            if (bit == 0) {
                const Frame next{f.t + 1, f.t, (f.prefix << 1) | 1u, 0};
                stack.push_back(next);
            }
        } else {
            stack.pop_back();
        }
        (void)L;
    }
}

template <class Ar>
using NecklaceKernel = void (*)(const uint64_t*, uint32_t, int, const int32_t*, int64_t, Ar, int64_t, uint32_t*,
                                uint64_t*);

template <class Ar, bool kCount = false>
NecklaceKernel<Ar> kernel_for(int L) {
    switch (L) {
#define PHIP_NECKLACE_CASE(X) \
    case X: return necklace_kernel<X, Ar, kCount>;
        PHIP_NECKLACE_CASE(2) PHIP_NECKLACE_CASE(4) PHIP_NECKLACE_CASE(6) PHIP_NECKLACE_CASE(10) PHIP_NECKLACE_CASE(12)
        PHIP_NECKLACE_CASE(16) PHIP_NECKLACE_CASE(18) PHIP_NECKLACE_CASE(22) PHIP_NECKLACE_CASE(28)
        PHIP_NECKLACE_CASE(30) PHIP_NECKLACE_CASE(36) PHIP_NECKLACE_CASE(40) PHIP_NECKLACE_CASE(42)
        PHIP_NECKLACE_CASE(46)
#undef PHIP_NECKLACE_CASE
        default: return nullptr;
    }
}

int smallest_primitive_root(int n) {
    std::vector<int> factors;
    int m = n - 1;
    for (int d = 2; d * d <= m; ++d) {
        if (m % d == 0) {
            factors.push_back(d);
            while (m % d == 0) m /= d;
        }
    }
    if (m > 1) factors.push_back(m);
    for (int g = 2; g < n; ++g) {
        bool ok = true;
        for (int f : factors) ok = ok && host_powmod((uint64_t)g, (uint64_t)((n - 1) / f), (uint64_t)n) != 1;
        if (ok) return g;
    }
    return -1;
}

}  // namespace

bool necklace_supported(int n, uint64_t p) {
    return n >= 5 && n <= kNecklaceMaxOrder && is_prime_u64((uint64_t)n) && (p & 1) && p >= 3 &&
           p < (1ull << 32) && kernel_for<NeckSigned31>(n - 1) != nullptr;
}

namespace {

// FKM task depth for word length L: at least L/2 (the kernel extends a
// prefix in one step), 14 levels above the leaves where the task list
// allows (fewer successor steps per representative in the slowest lane of
// a wave), at most the tuning limit (task list size).
int necklace_depth(int L) { return std::max(L / 2, std::min(L - 14, tuning().necklace_depth)); }

}  // namespace

int necklace_orbit_stats(DeviceContext* ctx, int n, uint64_t* representatives, uint64_t* words) {
    if (n < 3 || n > kNecklaceMaxOrder || !is_prime_u64((uint64_t)n) || !(n & 1)) {
        set_error("n must be an odd prime satisfying 3 <= n <= 47 on the device");
        return -1;
    }
    const int L = n - 1;
    const NecklaceKernel<NeckCount> kernel = kernel_for<NeckCount, true>(L);
    if (!kernel) {
        set_error("internal: no necklace kernel for this order");
        return -1;
    }
    const int depth = necklace_depth(L);
    std::vector<uint64_t> tasks;
    collect_tasks(L, depth, tasks);
    if (tasks.empty() || tasks.size() >= 0xffffffffull) {
        set_error("internal: necklace task list");
        return -1;
    }
    const int cus = ctx->compute_units() > 0 ? ctx->compute_units() : 256;
    int per_cu = 0;
    PHIP_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kNecklaceBlock, 0));
    const unsigned grid = (unsigned)(cus * std::max(1, per_cu));
    hipStream_t st = ctx->stream();
    auto* d_tasks = (uint64_t*)ctx->buffer(25, tasks.size() * 8);
    auto* d_E = (int32_t*)ctx->buffer(26, 2 * (size_t)L * 4);
    auto* d_cnt = (uint32_t*)ctx->buffer(27, 16);
    auto* d_part = (uint64_t*)ctx->buffer(28, (size_t)grid * 16);
    auto* h_part = (uint64_t*)ctx->host_buffer(5, (size_t)grid * 16);
    if (!d_tasks || !d_E || !d_cnt || !d_part || !h_part) return -1;
    const uint32_t zero = 0;
    PHIP_HIP_TRY(hipMemcpyAsync(d_tasks, tasks.data(), tasks.size() * 8, hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d_cnt, &zero, 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kNecklaceBlock), 0, st, d_tasks, (uint32_t)tasks.size(), depth, d_E,
                       (int64_t)0, NeckCount{}, (int64_t)1, d_cnt, d_part);
    PHIP_HIP_TRY(hipGetLastError());
    PHIP_HIP_TRY(hipMemcpyAsync(h_part, d_part, (size_t)grid * 16, hipMemcpyDeviceToHost, st));
    PHIP_HIP_TRY(hipStreamSynchronize(st));
    uint64_t reps = 0, covered = 0;
    for (unsigned b = 0; b < grid; ++b) {
        reps += h_part[b];
        covered += h_part[grid + b];
    }
    *representatives = reps;
    *words = covered;
    return 0;
}

int necklace_minor_modp(DeviceContext* ctx, int n, uint64_t omega, uint64_t p, int64_t* minor_out) {
    if (!necklace_supported(n, p)) {
        set_error("necklace path needs a prime order n <= 47 and an odd prime p < 2^32");
        return -1;
    }
    const int L = n - 1;
    const int g = smallest_primitive_root(n);
    if (g < 0) {
        set_error("internal: no primitive root");
        return -1;
    }
    // First row w[k] = omega^(g^k mod n), centred.
    std::vector<int64_t> first(L);
    int64_t base_sum = 0;
    uint64_t e = 1;
    for (int k = 0; k < L; ++k) {
        int64_t v = (int64_t)host_powmod(omega, e, p);
        if (v > (int64_t)(p / 2)) v -= (int64_t)p;
        first[k] = v;
        base_sum += v;
        e = e * (uint64_t)g % (uint64_t)n;
    }
    std::vector<int32_t> E(2 * L);
    for (int j = 0; j < 2 * L; ++j) E[j] = (int32_t)first[((-j) % L + L) % L];

    const int depth = necklace_depth(L);
    std::vector<uint64_t> tasks;
    collect_tasks(L, depth, tasks);
    if (tasks.empty() || tasks.size() >= 0xffffffffull) {
        set_error("internal: necklace task list");
        return -1;
    }
    const int cus = ctx->compute_units() > 0 ? ctx->compute_units() : 256;
    // As many workgroups per CU as the kernel's registers allow (3 at minor
    // order 40 and 46, 4 below): the waves hide each other's scalar-load
    // and cross-lane latency, and the task counter balances them.
    int per_cu = 0;
    if (p < (1ull << 31)) {
        PHIP_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel_for<NeckSigned31>(L),
                                                                  kNecklaceBlock, 0));
    } else {
        PHIP_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel_for<NeckUnsigned32>(L),
                                                                  kNecklaceBlock, 0));
    }
    const unsigned grid = (unsigned)(cus * std::max(1, per_cu));
    hipStream_t st = ctx->stream();
    auto* d_tasks = (uint64_t*)ctx->buffer(25, tasks.size() * 8);
    auto* d_E = (int32_t*)ctx->buffer(26, E.size() * 4);
    auto* d_cnt = (uint32_t*)ctx->buffer(27, 16);
    auto* d_part = (uint64_t*)ctx->buffer(28, (size_t)grid * 8);
    auto* h_part = (uint64_t*)ctx->host_buffer(5, (size_t)grid * 8);
    if (!d_tasks || !d_E || !d_cnt || !d_part || !h_part) return -1;
    const uint32_t zero = 0;
    PHIP_HIP_TRY(hipMemcpyAsync(d_tasks, tasks.data(), tasks.size() * 8, hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d_E, E.data(), E.size() * 4, hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d_cnt, &zero, 4, hipMemcpyHostToDevice, st));
    if (ctx->profiling) PHIP_HIP_TRY(hipEventRecord(ctx->prof_start, st));
    if (p < (1ull << 31)) {
        const NeckSigned31 ar{(int32_t)p, -(int32_t)p, inverse_mod_2_32((uint32_t)p)};
        hipLaunchKernelGGL(kernel_for<NeckSigned31>(L), dim3(grid), dim3(kNecklaceBlock), 0, st, d_tasks,
                           (uint32_t)tasks.size(), depth, d_E, base_sum, ar, (int64_t)p, d_cnt, d_part);
    } else {
        const NeckUnsigned32 ar{make_mont32(p), (int64_t)p << 5, (int64_t)p << 30};
        hipLaunchKernelGGL(kernel_for<NeckUnsigned32>(L), dim3(grid), dim3(kNecklaceBlock), 0, st, d_tasks,
                           (uint32_t)tasks.size(), depth, d_E, base_sum, ar, (int64_t)p, d_cnt, d_part);
    }
    PHIP_HIP_TRY(hipGetLastError());
    if (ctx->profiling) PHIP_HIP_TRY(hipEventRecord(ctx->prof_stop, st));
    PHIP_HIP_TRY(hipMemcpyAsync(h_part, d_part, (size_t)grid * 8, hipMemcpyDeviceToHost, st));
    PHIP_HIP_TRY(hipStreamSynchronize(st));
    uint64_t sum = 0;  // block sums are below 256 p < 2^40
    for (unsigned b = 0; b < grid; ++b) sum = (sum + h_part[b] % p) % p;
    if (ctx->profiling) {
        float ms = 0;
        PHIP_HIP_TRY(hipEventElapsedTime(&ms, ctx->prof_start, ctx->prof_stop));
        ctx->profiled_ms += ms;
        ctx->profiled_launches += 1;
    }
    // Remove R^-2L (L row reductions, L-1 tree products, one fold) and the
    // 2^L of Glynn's full sum.
    const uint64_t r_mod_p = (1ull << 32) % p;
    const uint64_t scale = host_mulmod(host_powmod(r_mod_p, (uint64_t)(2 * L), p), host_powmod((p + 1) / 2, (uint64_t)L, p), p);
    *minor_out = (int64_t)host_mulmod(sum, scale, p);
    return 0;
}

}  // namespace phip
