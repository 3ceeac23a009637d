// Shared declarations of the persistent batch (see batch.hip, batch_plan.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace phip {

// One (matrix, modulus) job of a batch plan; 48 bytes, read with scalar loads.
struct BatchJob {
    uint64_t data_off;  // int32 offset of the job's glynn_chunk tables (n * glynn_stride words)
    uint64_t lanes;     // lanes of this job (2^(n-1-cbits))
    int32_t n;
    int32_t cbits;
    int32_t p;
    uint32_t pinv;
    uint32_t first_item;
    uint32_t n_items;
    uint32_t scale;  // R^(groups-1) * 2^-(n-1) mod p
    int32_t group;
};

constexpr int kBatchMinOrder = 5;
constexpr int kBatchMaxOrder = 40;

using BatchLaunchFn = hipError_t (*)(dim3 grid, hipStream_t stream, const BatchJob* jobs, const uint2* items,
                                     uint32_t n_items, uint32_t* counter, const int32_t* blob,
                                     uint64_t* item_partials);

hipError_t launch_batch_g0(dim3, hipStream_t, const BatchJob*, const uint2*, uint32_t, uint32_t*, const int32_t*,
                           uint64_t*);
hipError_t launch_batch_g2(dim3, hipStream_t, const BatchJob*, const uint2*, uint32_t, uint32_t*, const int32_t*,
                           uint64_t*);
hipError_t launch_batch_g3(dim3, hipStream_t, const BatchJob*, const uint2*, uint32_t, uint32_t*, const int32_t*,
                           uint64_t*);
hipError_t launch_batch_g4(dim3, hipStream_t, const BatchJob*, const uint2*, uint32_t, uint32_t*, const int32_t*,
                           uint64_t*);
hipError_t launch_batch_g6(dim3, hipStream_t, const BatchJob*, const uint2*, uint32_t, uint32_t*, const int32_t*,
                           uint64_t*);

}  // namespace phip
