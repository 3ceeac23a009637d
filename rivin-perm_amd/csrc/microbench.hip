// Measurement helpers used by the benchmark to pin its roofline:
//
//   phip_valu_microbench  throughput of the VALU instructions the Glynn
//                         kernels issue (v_add_u32, v_mul_i32_i24,
//                         v_mul_lo_u32, v_mul_hi_i32, v_mad_i64_i32,
//                         v_mad_u64_u32, v_pk_add_u16, SDWA v_mul_i32_i24,
//                         v_lshl_add_u64, v_ashrrev_i32)
//                         and of one complete signed Montgomery product,
//                         each as lane-operations per
//                         second over the whole device.  Eight independent
//                         dependency chains per lane keep the pipeline full.
//   phip_mix_microbench   issue rate of a whole instruction mix (rounds of 8
//                         independent instructions per type and iteration),
//                         the roofline of a kernel's own instruction stream.
//   phip_copy_calibration a plain device-to-device copy kernel of a known
//                         size, to calibrate HBM byte counters.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "device_context.h"
#include "permhip.h"
#include "phip_arith.h"

namespace phip {

namespace {

constexpr int kChains = 8;
constexpr int kUnroll = 8;

__device__ __forceinline__ int32_t mont31_mb(int32_t a, int32_t b, int32_t negp, uint32_t pinv) {
    int64_t t = (int64_t)a * (int64_t)b;
    int32_t m = (int32_t)((uint32_t)t * pinv);
    int64_t u = t + (int64_t)m * (int64_t)negp;
    return (int32_t)(u >> 32);
}

template <int OP>
__global__ __launch_bounds__(256) void valu_rate_kernel(uint32_t* out, int iters, uint32_t c) {
    uint32_t x[kChains];
    uint64_t y[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) {
        x[k] = threadIdx.x * 2654435761u + k + c;
        y[k] = x[k];
    }
    const int32_t p = 2147483647;
    const uint32_t pinv = inverse_mod_2_32((uint32_t)p);
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
            for (int k = 0; k < kChains; ++k) {
                if constexpr (OP == 0) {
                    asm volatile("v_add_u32 %0, %0, %1" : "+v"(x[k]) : "v"(c));
                } else if constexpr (OP == 1) {
                    asm volatile("v_mul_i32_i24 %0, %0, %1" : "+v"(x[k]) : "v"(c));
                } else if constexpr (OP == 2) {
                    asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(x[k]) : "v"(c));
                } else if constexpr (OP == 3) {
                    asm volatile("v_mul_hi_i32 %0, %0, %1" : "+v"(x[k]) : "v"(c));
                } else if constexpr (OP == 4) {
                    asm volatile("v_mad_i64_i32 %0, vcc, %1, %2, %0" : "+v"(y[k]) : "v"(x[k]), "v"(c) : "vcc");
                } else if constexpr (OP == 5) {
                    asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(y[k]) : "v"(x[k]), "v"(c) : "vcc");
                } else if constexpr (OP == 6) {
                    x[k] = (uint32_t)mont31_mb((int32_t)x[k], (int32_t)c, -p, pinv);
                } else if constexpr (OP == 7) {
                    asm volatile("v_pk_add_u16 %0, %0, %1" : "+v"(x[k]) : "v"(c));
                } else if constexpr (OP == 9) {
                    asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(y[k]) : "v"((uint64_t)c));
                } else if constexpr (OP == 10) {
                    asm volatile("v_ashrrev_i32 %0, 1, %0" : "+v"(x[k]));
                } else {
                    asm volatile(
                        "v_mul_i32_i24_sdwa %0, sext(%0), sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD "
                        "src0_sel:WORD_1 src1_sel:WORD_0"
                        : "+v"(x[k])
                        : "v"(c));
                }
            }
        }
    }
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < kChains; ++k) acc ^= x[k] ^ (uint32_t)y[k] ^ (uint32_t)(y[k] >> 32);
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

// Issue bound of an instruction mix: every iteration issues rounds[t]
// rounds of 8 independent instructions of type t (v_mad_i64_i32,
// v_mul_lo_u32, v_pk_add_u16, SDWA v_mul_i32_i24, v_add_u32), each type on
// its own chains so nothing but issue limits the rate.
__global__ __launch_bounds__(256) void mix_kernel(uint32_t* out, int iters, int r_mad, int r_mul, int r_pk,
                                                  int r_m24, int r_add, uint32_t c) {
    uint64_t y[kChains];
    uint32_t a[kChains], b[kChains], d[kChains], e[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) {
        y[k] = threadIdx.x + k;
        a[k] = b[k] = d[k] = e[k] = threadIdx.x * 2654435761u + k + c;
    }
    for (int it = 0; it < iters; ++it) {
        for (int r = 0; r < r_mad; ++r) {
#pragma unroll
            for (int k = 0; k < kChains; ++k) {
                uint64_t carry;
                asm volatile("v_mad_i64_i32 %0, %1, %2, %3, %0" : "+v"(y[k]), "=s"(carry) : "v"(c + k), "v"(c));
            }
        }
        for (int r = 0; r < r_mul; ++r) {
#pragma unroll
            for (int k = 0; k < kChains; ++k) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(a[k]) : "v"(c));
        }
        for (int r = 0; r < r_pk; ++r) {
#pragma unroll
            for (int k = 0; k < kChains; ++k) asm volatile("v_pk_add_u16 %0, %0, %1" : "+v"(b[k]) : "v"(c));
        }
        for (int r = 0; r < r_m24; ++r) {
#pragma unroll
            for (int k = 0; k < kChains; ++k)
                asm volatile(
                    "v_mul_i32_i24_sdwa %0, sext(%0), sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD "
                    "src0_sel:WORD_1 src1_sel:WORD_0"
                    : "+v"(d[k])
                    : "v"(c));
        }
        for (int r = 0; r < r_add; ++r) {
#pragma unroll
            for (int k = 0; k < kChains; ++k) asm volatile("v_add_u32 %0, %0, %1" : "+v"(e[k]) : "v"(c));
        }
    }
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < kChains; ++k) acc ^= (uint32_t)y[k] ^ (uint32_t)(y[k] >> 32) ^ a[k] ^ b[k] ^ d[k] ^ e[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                   uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        dst[i] = src[i];
}

}  // namespace

}  // namespace phip

extern "C" {

int phip_valu_microbench(double* lane_ops_per_s, int len) {
    phip::DeviceContext* ctx = phip::DeviceContext::current();
    if (!ctx) return -1;
    std::lock_guard<std::mutex> lock(ctx->mutex());
    const int cus = ctx->compute_units() > 0 ? ctx->compute_units() : 256;
    const int blocks = cus * 8;  // 32 waves per CU
    const int iters = 4096;
    hipStream_t st = ctx->stream();
    auto* out = (uint32_t*)ctx->buffer(0, (size_t)blocks * 256 * 4);
    if (!out) return -1;
    hipEvent_t e0, e1;
    PHIP_HIP_TRY(hipEventCreate(&e0));
    PHIP_HIP_TRY(hipEventCreate(&e1));
    using Kernel = void (*)(uint32_t*, int, uint32_t);
    const Kernel kernels[] = {phip::valu_rate_kernel<0>, phip::valu_rate_kernel<1>, phip::valu_rate_kernel<2>,
                              phip::valu_rate_kernel<3>, phip::valu_rate_kernel<4>, phip::valu_rate_kernel<5>,
                              phip::valu_rate_kernel<6>, phip::valu_rate_kernel<7>, phip::valu_rate_kernel<8>,
                              phip::valu_rate_kernel<9>, phip::valu_rate_kernel<10>};
    const int nk = (int)(sizeof(kernels) / sizeof(kernels[0]));
    int written = 0;
    for (int k = 0; k < nk && k < len; ++k) {
        // Warm-up, then one timed launch.
        hipLaunchKernelGGL(kernels[k], dim3(blocks), dim3(256), 0, st, out, 64, 12345u);
        PHIP_HIP_TRY(hipGetLastError());
        PHIP_HIP_TRY(hipEventRecord(e0, st));
        hipLaunchKernelGGL(kernels[k], dim3(blocks), dim3(256), 0, st, out, iters, 12345u);
        PHIP_HIP_TRY(hipGetLastError());
        PHIP_HIP_TRY(hipEventRecord(e1, st));
        PHIP_HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        PHIP_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        const double ops = (double)blocks * 256 * iters * phip::kUnroll * phip::kChains;
        lane_ops_per_s[k] = ops / (ms * 1e-3);
        ++written;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return written;
}

int phip_mix_microbench(const int32_t* rounds, int iters, double* lane_iters_per_s) {
    if (!rounds || !lane_iters_per_s || iters < 1 || iters > (1 << 20)) {
        phip::set_error("invalid mix microbenchmark arguments");
        return -1;
    }
    for (int t = 0; t < 5; ++t) {
        if (rounds[t] < 0 || rounds[t] > 4096) {
            phip::set_error("mix rounds must lie in [0, 4096]");
            return -1;
        }
    }
    phip::DeviceContext* ctx = phip::DeviceContext::current();
    if (!ctx) return -1;
    std::lock_guard<std::mutex> lock(ctx->mutex());
    const int cus = ctx->compute_units() > 0 ? ctx->compute_units() : 256;
    const int blocks = cus * 8;  // 32 waves per CU
    hipStream_t st = ctx->stream();
    auto* out = (uint32_t*)ctx->buffer(0, (size_t)blocks * 256 * 4);
    if (!out) return -1;
    hipEvent_t e0, e1;
    PHIP_HIP_TRY(hipEventCreate(&e0));
    PHIP_HIP_TRY(hipEventCreate(&e1));
    hipLaunchKernelGGL(phip::mix_kernel, dim3(blocks), dim3(256), 0, st, out, 1, rounds[0], rounds[1], rounds[2],
                       rounds[3], rounds[4], 12345u);
    PHIP_HIP_TRY(hipGetLastError());
    PHIP_HIP_TRY(hipEventRecord(e0, st));
    hipLaunchKernelGGL(phip::mix_kernel, dim3(blocks), dim3(256), 0, st, out, iters, rounds[0], rounds[1],
                       rounds[2], rounds[3], rounds[4], 12345u);
    PHIP_HIP_TRY(hipGetLastError());
    PHIP_HIP_TRY(hipEventRecord(e1, st));
    PHIP_HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    PHIP_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *lane_iters_per_s = (double)blocks * 256 * iters / (ms * 1e-3);
    return 0;
}

int phip_copy_calibration(uint64_t bytes, int reps) {
    phip::DeviceContext* ctx = phip::DeviceContext::current();
    if (!ctx) return -1;
    std::lock_guard<std::mutex> lock(ctx->mutex());
    const uint64_t n = (bytes + 15) / 16;
    auto* src = (uint4*)ctx->buffer(0, n * 16);
    auto* dst = (uint4*)ctx->buffer(1, n * 16);
    if (!src || !dst) return -1;
    hipStream_t st = ctx->stream();
    for (int r = 0; r < reps; ++r) {
        hipLaunchKernelGGL(phip::copy_kernel, dim3(2048), dim3(256), 0, st, src, dst, n);
        PHIP_HIP_TRY(hipGetLastError());
    }
    PHIP_HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"
