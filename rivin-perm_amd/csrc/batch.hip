// Persistent, work-stealing Glynn kernel for 31-bit odd moduli (gfx950).
//
// A plan holds jobs of uneven order; each job is one table with a set of K
// moduli that share it.  Every job is cut into items of one wavefront each
// (64 lanes x 2^cbits Gray states, see gray_walk).  Items are sorted by cost,
// largest first, and grouped by row-sum representation G and set size K;
// each (G, K) class runs as one persistent launch whose wavefronts take
// items from a global counter until the list is exhausted.  An item's
// wavefront sums (one per prime of the set; signed 64-bit sums of the lanes'
// Montgomery residues) go to its own slots, and a finalisation kernel folds
// each job's contiguous slots into residues, so no atomics touch the
// arithmetic.
//
// This file is compiled once per (G, K) class: PHIP_BATCH_GROUP and
// PHIP_BATCH_SET select it, PHIP_BATCH_PAIRED=1 the variant for prime
// sets below kPairedPrimeLimit that pairs consecutive states, and
// PHIP_BATCH_SKIP_ZEROS=1 the variant for tables whose states are mostly
// zero terms (ZeroSkipSink), whose item sums cover every lane, and
// PHIP_BATCH_EXACT=1 the exact-accumulation variant (ExactSink), whose items
// hold PHIP_BATCH_SET limb columns of the exact sums for all of a job's
// primes, and PHIP_BATCH_COMPACT=1 (with it) the form whose product tree is
// sized for kCompactValueBits-bit group values (LimbProduct), and
// PHIP_BATCH_SIGNED=1 (with it, PHIP_BATCH_SET=8) the signed-limb form
// (slimb_product8).  Both
// PHIP_BATCH_SKIP_ZEROS=1 and PHIP_BATCH_EXACT=1 (with PHIP_BATCH_SET=8)
// select the block zero-skipping walk with exact accumulation
// (glynn_chunk_blocks_exact): 0/1 block tables of at most eight groups.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "batch.h"
#include "glynn31.h"

#ifndef PHIP_BATCH_GROUP
#error "PHIP_BATCH_GROUP must be defined"
#endif
#ifndef PHIP_BATCH_SET
#error "PHIP_BATCH_SET must be defined"
#endif
#ifndef PHIP_BATCH_PAIRED
#define PHIP_BATCH_PAIRED 0
#endif
#ifndef PHIP_BATCH_SKIP_ZEROS
#define PHIP_BATCH_SKIP_ZEROS 0
#endif
#ifndef PHIP_BATCH_EXACT
#define PHIP_BATCH_EXACT 0
#endif
#ifndef PHIP_BATCH_COMPACT
#define PHIP_BATCH_COMPACT 0
#endif
#ifndef PHIP_BATCH_SIGNED
#define PHIP_BATCH_SIGNED 0
#endif
#if PHIP_BATCH_SIGNED && (!PHIP_BATCH_EXACT || PHIP_BATCH_COMPACT || PHIP_BATCH_SKIP_ZEROS || PHIP_BATCH_SET != 8)
#error "the signed-limb classes are 8-column exact-accumulation classes"
#endif
#if PHIP_BATCH_COMPACT && !PHIP_BATCH_EXACT
#error "the compact classes are exact-accumulation classes"
#endif
// Width of the values the exact classes' product trees are sized for.
constexpr int kBatchValueBits =
    PHIP_BATCH_SIGNED ? phip::kSignedLimbs : (PHIP_BATCH_COMPACT ? phip::kCompactValueBits : 31);

namespace phip {

namespace {

// One wavefront's chunk of job `jb` for every prime of its set; returns the
// lane values (signed Montgomery residues) in r.
template <int N, int G, int K, bool P, bool Z>
__device__ __forceinline__ void run_item(const int32_t* __restrict__ tables, const BatchJob& jb, uint64_t base,
                                         bool active, int32_t* queue, bool reload, const int32_t (&p)[K],
                                         const uint32_t (&pinv)[K], int32_t (&r)[K]) {
    if constexpr (Z) {
        if constexpr (N >= kBlockMinOrder) {
            if (jb.block_rows >= 0) {
                // Modulo-3 jobs have one modulus, padded to a set of 2.
                if constexpr (K == 2) {
                    if (jb.block_mod3) {
                        glynn_chunk_set_blocks<N, G, K, true>(tables, N, jb.cbits, base, active, queue,
                                                              jb.block_rows, reload, p, pinv, r);
                        return;
                    }
                }
                glynn_chunk_set_blocks<N, G, K>(tables, N, jb.cbits, base, active, queue, jb.block_rows, reload, p,
                                                pinv, r);
                return;
            }
        }
        if constexpr (N >= kSkipHalfMinOrder) {
            if (jb.test_words < skip_full_words(N)) {
                glynn_chunk_set_skip<N, G, K, true>(tables, N, jb.cbits, base, active, queue, p, pinv, r);
                return;
            }
        }
        glynn_chunk_set_skip<N, G, K, false>(tables, N, jb.cbits, base, active, queue, p, pinv, r);
    } else if constexpr (K == 1) {
        r[0] = glynn_chunk<N, G>(tables, N, jb.cbits, base, p[0], pinv[0]);
    } else {
        glynn_chunk_set<N, G, K, P>(tables, N, jb.cbits, base, p, pinv, r);
    }
}

// One wavefront's chunk of an exact job: the lanes' exact signed sums as L
// limb columns each.  The L = 8 class takes tables of at most
// kExactMaxGroups groups, the wider ones tables of more groups, up to
// kExactWideMaxGroups; other orders are never planned on the class and
// compile to nothing.
template <int N, int G, int L, int V, bool Z>
__device__ __forceinline__ void run_item_exact(const int32_t* __restrict__ tables, const BatchJob& jb, uint64_t base,
                                               bool active, int32_t* queue, bool reload, int64_t (&cols)[L]) {
    constexpr int kGroups = G > 0 ? (N + G - 1) / G : 0;
    if constexpr (Z) {
        // Block tables of at most eight groups (the planner sends no others).
        if constexpr (N >= kBlockMinOrder && L == kExactColumns && G > 0 && kGroups <= kExactMaxGroups) {
            glynn_chunk_blocks_exact<N, G>(tables, N, jb.cbits, base, active, queue, jb.block_rows, reload, cols);
        }
    } else if constexpr (G > 0 && (L == kExactColumns ? kGroups <= kExactMaxGroups
                                                      : (kGroups > kExactMaxGroups && kGroups <= kExactWideMaxGroups))) {
        glynn_chunk_exact<N, G, L, V>(tables, N, jb.cbits, base, cols);
    }
}

static_assert(kZeroQueueEntries == kZeroQueue, "zero-skip queue sizes disagree");

template <int G, int K, bool P, bool Z, bool E, int V>
__global__ __launch_bounds__(256, 2) void batch_glynn_kernel(const BatchJob* __restrict__ jobs,
                                                         const uint2* __restrict__ items, uint32_t n_items,
                                                         uint32_t* __restrict__ counter,
                                                         const int32_t* __restrict__ blob,
                                                         const int32_t* __restrict__ set_p,
                                                         const uint32_t* __restrict__ set_pinv,
                                                         uint64_t* __restrict__ item_partials,
                                                         uint32_t queue_words) {
    const int lane_id = threadIdx.x & 63;
    const uint32_t claim = lane_id == 0 ? 1u : 0u;
    // One zero-skip queue of queue_words words per wavefront (4 per
    // workgroup) in dynamic LDS: skip_queue_rows(n) * kZeroQueue words for
    // the class's largest order n, set by the host.
    extern __shared__ int32_t queues[];
    int32_t* const queue = queues + (Z ? (threadIdx.x >> 6) * queue_words : 0);
    // The job whose block-walk tables this wavefront's LDS region holds.
    uint32_t loaded_job = 0xffffffffu;
    for (;;) {
        // Every lane executes the same atomic (lane 0 adds 1, the others 0)
        // so no lane-dependent branch sits in front of the broadcast: with a
        // lane-0-only branch here the compiler split the loop so that lanes
        // 1..63 re-read a stale index without lane 0 and never left it.
        const uint32_t old = atomicAdd(counter, claim);
        const uint32_t idx = (uint32_t)__builtin_amdgcn_readfirstlane((int)old);
        if (idx >= n_items) break;
        const uint2 it = items[idx];
        const BatchJob jb = jobs[it.x];
        const uint64_t lane = (uint64_t)it.y * 64 + lane_id;
        const bool active = lane < jb.lanes;
        const uint64_t base = (active ? lane : 0) << jb.cbits;
        const int32_t* __restrict__ tables = blob + jb.data_off;
        // A block-walk wavefront keeps its job's tables in LDS between items.
        const bool block = Z && jb.block_rows >= 0;
        const bool reload = block && it.x != loaded_job;
        loaded_job = block ? it.x : 0xffffffffu;
        if constexpr (E) {
            // Exact jobs: K limb columns per item, for all of the job's
            // primes (the finalisation reduces them per prime).
            int64_t cols[K];
#pragma unroll
            for (int c = 0; c < K; ++c) cols[c] = 0;
            switch (jb.n) {
#define PHIP_EXACT_CASE(N) \
    case N:                \
        run_item_exact<N, G, K, V, Z>(tables, jb, base, active, queue, reload, cols); \
        break;
                PHIP_EXACT_CASE(5) PHIP_EXACT_CASE(6) PHIP_EXACT_CASE(7) PHIP_EXACT_CASE(8)
                PHIP_EXACT_CASE(9) PHIP_EXACT_CASE(10) PHIP_EXACT_CASE(11) PHIP_EXACT_CASE(12)
                PHIP_EXACT_CASE(13) PHIP_EXACT_CASE(14) PHIP_EXACT_CASE(15) PHIP_EXACT_CASE(16)
                PHIP_EXACT_CASE(17) PHIP_EXACT_CASE(18) PHIP_EXACT_CASE(19) PHIP_EXACT_CASE(20)
                PHIP_EXACT_CASE(21) PHIP_EXACT_CASE(22) PHIP_EXACT_CASE(23) PHIP_EXACT_CASE(24)
                PHIP_EXACT_CASE(25) PHIP_EXACT_CASE(26) PHIP_EXACT_CASE(27) PHIP_EXACT_CASE(28)
                PHIP_EXACT_CASE(29) PHIP_EXACT_CASE(30) PHIP_EXACT_CASE(31) PHIP_EXACT_CASE(32)
                PHIP_EXACT_CASE(33) PHIP_EXACT_CASE(34) PHIP_EXACT_CASE(35) PHIP_EXACT_CASE(36)
                PHIP_EXACT_CASE(37) PHIP_EXACT_CASE(38) PHIP_EXACT_CASE(39) PHIP_EXACT_CASE(40)
                PHIP_EXACT_CASE(41) PHIP_EXACT_CASE(42) PHIP_EXACT_CASE(43) PHIP_EXACT_CASE(44)
                PHIP_EXACT_CASE(45) PHIP_EXACT_CASE(46)
#undef PHIP_EXACT_CASE
                default:
                    break;
            }
#pragma unroll
            for (int c = 0; c < K; ++c) {
                // Column sums of 64 lanes in two's complement: |sum| <
                // 2^(37 + cbits).  Block-walk lanes take states of any lane.
                const uint64_t s = wave_reduce_sum_u64((Z || active) ? (uint64_t)cols[c] : 0u);
                if (lane_id == 0) item_partials[(size_t)idx * kItemSlots + c] = s;
            }
        } else {
        int32_t p[K];
        uint32_t pinv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            p[k] = __builtin_amdgcn_readfirstlane(set_p[jb.prime_off + k]);
            pinv[k] = (uint32_t)__builtin_amdgcn_readfirstlane((int)set_pinv[jb.prime_off + k]);
        }
        int32_t r[K];
#pragma unroll
        for (int k = 0; k < K; ++k) r[k] = 0;
        switch (jb.n) {
#define PHIP_BATCH_CASE(N) \
    case N:                \
        run_item<N, G, K, P, Z>(tables, jb, base, active, queue, reload, p, pinv, r); \
        break;
            PHIP_BATCH_CASE(5) PHIP_BATCH_CASE(6) PHIP_BATCH_CASE(7) PHIP_BATCH_CASE(8)
            PHIP_BATCH_CASE(9) PHIP_BATCH_CASE(10) PHIP_BATCH_CASE(11) PHIP_BATCH_CASE(12)
            PHIP_BATCH_CASE(13) PHIP_BATCH_CASE(14) PHIP_BATCH_CASE(15) PHIP_BATCH_CASE(16)
            PHIP_BATCH_CASE(17) PHIP_BATCH_CASE(18) PHIP_BATCH_CASE(19) PHIP_BATCH_CASE(20)
            PHIP_BATCH_CASE(21) PHIP_BATCH_CASE(22) PHIP_BATCH_CASE(23) PHIP_BATCH_CASE(24)
            PHIP_BATCH_CASE(25) PHIP_BATCH_CASE(26) PHIP_BATCH_CASE(27) PHIP_BATCH_CASE(28)
            PHIP_BATCH_CASE(29) PHIP_BATCH_CASE(30) PHIP_BATCH_CASE(31) PHIP_BATCH_CASE(32)
            PHIP_BATCH_CASE(33) PHIP_BATCH_CASE(34) PHIP_BATCH_CASE(35) PHIP_BATCH_CASE(36)
            PHIP_BATCH_CASE(37) PHIP_BATCH_CASE(38) PHIP_BATCH_CASE(39) PHIP_BATCH_CASE(40)
            PHIP_BATCH_CASE(41) PHIP_BATCH_CASE(42) PHIP_BATCH_CASE(43) PHIP_BATCH_CASE(44)
            PHIP_BATCH_CASE(45) PHIP_BATCH_CASE(46)
#undef PHIP_BATCH_CASE
            default:
                break;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            // Signed lane values summed in two's complement: |sum| < 2^37.
            // Zero-skip lanes hold terms of any lane's chunk (inactive
            // lanes queue nothing but may take queued terms).
            const uint64_t s = wave_reduce_sum_u64((Z || active) ? (uint64_t)(int64_t)r[k] : 0u);
            if (lane_id == 0) item_partials[(size_t)idx * kItemSlots + k] = s;
        }
        }
    }
}

}  // namespace

#define PHIP_CAT5(a, b, c, d, e) a##b##c##d##e
#define PHIP_LAUNCH_NAME(G, K, P) PHIP_CAT5(launch_batch_g, G, _k, K, P)
#if PHIP_BATCH_PAIRED && (PHIP_BATCH_SKIP_ZEROS + PHIP_BATCH_EXACT) > 0
#error "a paired class neither skips zero terms nor accumulates exactly"
#elif PHIP_BATCH_SKIP_ZEROS && PHIP_BATCH_EXACT
#if PHIP_BATCH_SET != 8 || PHIP_BATCH_COMPACT
#error "exact block walks hold 8 limb columns per item"
#endif
#define PHIP_PAIRED_SUFFIX _skipexact
#elif PHIP_BATCH_EXACT
#if PHIP_BATCH_SET != 8 && PHIP_BATCH_SET != 10 && PHIP_BATCH_SET != 12
#error "exact classes hold 8, 10 or 12 limb columns per item"
#endif
#if PHIP_BATCH_COMPACT
#define PHIP_PAIRED_SUFFIX _compact
#elif PHIP_BATCH_SIGNED
#define PHIP_PAIRED_SUFFIX _signed
#else
#define PHIP_PAIRED_SUFFIX _exact
#endif
#elif PHIP_BATCH_PAIRED
#define PHIP_PAIRED_SUFFIX _paired
#elif PHIP_BATCH_SKIP_ZEROS
#define PHIP_PAIRED_SUFFIX _skip
#else
#define PHIP_PAIRED_SUFFIX
#endif

hipError_t PHIP_LAUNCH_NAME(PHIP_BATCH_GROUP, PHIP_BATCH_SET, PHIP_PAIRED_SUFFIX)(dim3 grid, unsigned lds_bytes,
                                                              hipStream_t stream, const BatchJob* jobs,
                                                              const uint2* items, uint32_t n_items,
                                                              uint32_t* counter, const int32_t* blob,
                                                              const int32_t* set_p, const uint32_t* set_pinv,
                                                              uint64_t* item_partials) {
    hipLaunchKernelGGL((batch_glynn_kernel<PHIP_BATCH_GROUP, PHIP_BATCH_SET, PHIP_BATCH_PAIRED != 0,
                                           PHIP_BATCH_SKIP_ZEROS != 0, PHIP_BATCH_EXACT != 0, kBatchValueBits>),
                       grid, dim3(256), lds_bytes, stream, jobs,
                       items, n_items, counter, blob, set_p, set_pinv, item_partials, lds_bytes / 16u);
    return hipGetLastError();
}

#define PHIP_OCC_NAME(G, K, P) PHIP_CAT5(occupancy_batch_g, G, _k, K, P)

// Workgroups of this class's kernel that one CU holds at once (its register
// and LDS use), or -1.
int PHIP_OCC_NAME(PHIP_BATCH_GROUP, PHIP_BATCH_SET, PHIP_PAIRED_SUFFIX)(unsigned lds_bytes) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &blocks,
            batch_glynn_kernel<PHIP_BATCH_GROUP, PHIP_BATCH_SET, PHIP_BATCH_PAIRED != 0, PHIP_BATCH_SKIP_ZEROS != 0,
                               PHIP_BATCH_EXACT != 0, kBatchValueBits>,
            256, lds_bytes) != hipSuccess)
        return -1;
    return blocks;
}

}  // namespace phip
