// Persistent, work-stealing batch of exact-permanent residues (gfx950).
//
// A plan holds many matrices of uneven order, each with its own list of odd
// moduli below 2^31.  Every (matrix, modulus) pair is a job; every job is cut
// into items of one wavefront each (64 lanes x 2^cbits Gray states, the same
// lane chunk as glynn31_kernel).  Items are sorted by cost, largest first,
// and grouped by row-sum representation (group size G); each group runs as
// one persistent launch whose wavefronts take items from a global counter
// until the list is exhausted.  Each item's wavefront sum goes to its own
// slot, and a finalisation kernel folds each job's contiguous slots into a
// residue, so no atomics touch the arithmetic.  The whole sequence (counter
// reset, persistent launches, finalisation, residue copy) is captured once
// into a hipGraph and replayed.
//
// The matrices and tables stay resident in device memory for the lifetime
// of the plan, so a replay moves only the residues back to the host.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "batch.h"
#include "glynn31.h"

#ifndef PHIP_BATCH_GROUP
#error "PHIP_BATCH_GROUP must be defined"
#endif

namespace phip {

#define PHIP_CAT2(a, b) a##b
#define PHIP_CAT(a, b) PHIP_CAT2(a, b)

namespace {

template <int G>
__global__ __launch_bounds__(256) void batch_glynn_kernel(const BatchJob* __restrict__ jobs,
                                                         const uint2* __restrict__ items, uint32_t n_items,
                                                         uint32_t* __restrict__ counter,
                                                         const int32_t* __restrict__ blob,
                                                         uint64_t* __restrict__ item_partials) {
    const int lane_id = threadIdx.x & 63;
    const uint32_t claim = lane_id == 0 ? 1u : 0u;
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
        uint64_t r = 0;
        switch (jb.n) {
#define PHIP_BATCH_CASE(N) \
    case N:                \
        r = glynn_chunk<N, G>(tables, N, jb.cbits, base, jb.p, jb.pinv); \
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
#undef PHIP_BATCH_CASE
            default:
                break;
        }
        const uint64_t s = wave_reduce_sum_u64(active ? r : 0u);
        if (lane_id == 0) item_partials[idx] = s;
    }
}

}  // namespace

hipError_t PHIP_CAT(launch_batch_g, PHIP_BATCH_GROUP)(dim3 grid, hipStream_t stream, const BatchJob* jobs,
                                                      const uint2* items, uint32_t n_items, uint32_t* counter,
                                                      const int32_t* blob, uint64_t* item_partials) {
    hipLaunchKernelGGL(batch_glynn_kernel<PHIP_BATCH_GROUP>, grid, dim3(256), 0, stream, jobs, items, n_items,
                       counter, blob, item_partials);
    return hipGetLastError();
}

}  // namespace phip
