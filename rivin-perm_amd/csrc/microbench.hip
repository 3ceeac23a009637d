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
//                         independent instructions per type and iteration,
//                         interleaved evenly), the roofline of a Glynn
//                         kernel's own instruction stream.
//   phip_copy_calibration a plain device-to-device copy kernel of a known
//                         size, to calibrate HBM byte counters.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

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

// Issue bound of an instruction mix: every iteration issues C_t rounds of
// 8 independent instructions of type t (v_mad_i64_i32, v_mul_lo_u32,
// v_pk_add_u16, SDWA v_mul_i32_i24, v_add_u32), each type on its own
// chains with register operands, interleaved evenly over R = max C_t steps
// (type t issues in step r when floor((r+1) C_t / R) > floor(r C_t / R)),
// as the instructions of one Gray state interleave in the Glynn kernels.
// The interleaving is fixed at compile time, so only the mixes of the
// compiled Glynn classes the benchmark reports are instantiated.
typedef uint16_t mb_u16x2 __attribute__((ext_vector_type(2)));

struct MixRegs {
    int64_t y[kChains];
    int32_t x[kChains];
    uint32_t a[4][kChains];
};

// One round of type T on 8 chains.  Plain arithmetic lets the compiler emit
// the instructions (and none of the hazard padding it adds around inline
// asm); `c` is re-hidden every step, so nothing folds.
template <int T>
__device__ __forceinline__ void mix_issue(MixRegs& m, uint32_t c) {
#pragma unroll
    for (int k = 0; k < kChains; ++k) {
        if constexpr (T == 0) {
            m.y[k] += (int64_t)m.x[k] * (int64_t)(int32_t)c;
        } else if constexpr (T == 1) {
            m.a[0][k] *= c;
        } else if constexpr (T == 2) {
            m.a[1][k] = __builtin_bit_cast(uint32_t, __builtin_bit_cast(mb_u16x2, m.a[1][k]) +
                                                         __builtin_bit_cast(mb_u16x2, c));
        } else if constexpr (T == 3) {
            asm volatile(
                "v_mul_i32_i24_sdwa %0, sext(%0), sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD "
                "src0_sel:WORD_1 src1_sel:WORD_0"
                : "+v"(m.a[2][k])
                : "v"(c));
        } else {
            m.a[3][k] += c;
        }
    }
}

constexpr bool mix_step(int count, int R, int r) { return (r + 1) * count / R > r * count / R; }

template <int C0, int C1, int C2, int C3, int C4, int R, int r>
__device__ __forceinline__ void mix_rounds(MixRegs& m, uint32_t& c) {
    if constexpr (r < R) {
        asm volatile("" : "+v"(c));
        if constexpr (mix_step(C0, R, r)) mix_issue<0>(m, c);
        if constexpr (mix_step(C1, R, r)) mix_issue<1>(m, c);
        if constexpr (mix_step(C2, R, r)) mix_issue<2>(m, c);
        if constexpr (mix_step(C3, R, r)) mix_issue<3>(m, c);
        if constexpr (mix_step(C4, R, r)) mix_issue<4>(m, c);
        mix_rounds<C0, C1, C2, C3, C4, R, r + 1>(m, c);
    }
}

constexpr int mix_max(int a, int b) { return a > b ? a : b; }

template <int C0, int C1, int C2, int C3, int C4>
__global__ __launch_bounds__(256) void mix_kernel(uint32_t* out, int iters, uint32_t c0) {
    constexpr int R = mix_max(mix_max(mix_max(C0, C1), mix_max(C2, C3)), C4);
    MixRegs m;
#pragma unroll
    for (int k = 0; k < kChains; ++k) {
        m.y[k] = threadIdx.x + k;
        m.x[k] = (int32_t)(threadIdx.x * 2654435761u + k + c0);
        for (int t = 0; t < 4; ++t) m.a[t][k] = (uint32_t)m.x[k] + t;
    }
    uint32_t c = c0 + threadIdx.x;
    for (int it = 0; it < iters; ++it) mix_rounds<C0, C1, C2, C3, C4, R, 0>(m, c);
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < kChains; ++k)
        acc ^= (uint32_t)m.y[k] ^ (uint32_t)((uint64_t)m.y[k] >> 32) ^ m.a[0][k] ^ m.a[1][k] ^ m.a[2][k] ^ m.a[3][k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

using MixKernel = void (*)(uint32_t*, int, uint32_t);

// Mixes of the compiled Glynn classes, per Gray state of n = 32 with groups
// of 4: the 8-prime set (per state, and with consecutive states paired) and
// the single prime.
struct MixEntry {
    int32_t rounds[5];
    MixKernel kernel;
};
const MixEntry kMixes[] = {
    {{92, 56, 16, 24, 2}, mix_kernel<92, 56, 16, 24, 2>},
    {{84, 52, 16, 24, 2}, mix_kernel<84, 52, 16, 24, 2>},
    {{15, 7, 16, 24, 2}, mix_kernel<15, 7, 16, 24, 2>},
};

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
    phip::MixKernel kernel = nullptr;
    for (const auto& e : phip::kMixes) {
        if (memcmp(e.rounds, rounds, sizeof(e.rounds)) == 0) kernel = e.kernel;
    }
    if (!kernel) {
        phip::set_error("no compiled instruction mix matches these rounds");
        return -1;
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
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), 0, st, out, 1, 12345u);
    PHIP_HIP_TRY(hipGetLastError());
    PHIP_HIP_TRY(hipEventRecord(e0, st));
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), 0, st, out, iters, 12345u);
    PHIP_HIP_TRY(hipGetLastError());
    PHIP_HIP_TRY(hipEventRecord(e1, st));
    PHIP_HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    PHIP_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
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
