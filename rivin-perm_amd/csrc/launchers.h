// Host-side launch helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace phip {

// Orders served by the register-resident 31-bit Glynn kernels.  Smaller
// orders (too few states for a 16-state unrolled chunk) and larger ones use
// the generic kernels.
constexpr int kGlynnMinOrder = 5;
constexpr int kGlynnMaxOrder = 40;
constexpr int kGlynnBlock = 256;  // threads per workgroup of glynn31_kernel

// Words per table column: residues one per word (group 0) or exact row sums
// packed two 16-bit halves per word (groups 2..4).
__host__ __device__ constexpr int glynn_stride(int n, int group) { return group == 0 ? n : (n + 1) / 2; }

struct Glynn31Args {
    const int32_t* tables;  // per prime: n * glynn_stride words (state-0 row sums, then n-1 delta columns)
    const int32_t* primes;
    const uint32_t* pinvs;
    int n;
    int cbits;
    uint64_t lanes_per_prime;
    uint64_t* partials;
};

hipError_t launch_glynn31_g0(int nr, dim3 grid, hipStream_t stream, const Glynn31Args& a);
hipError_t launch_glynn31_g2(int nr, dim3 grid, hipStream_t stream, const Glynn31Args& a);
hipError_t launch_glynn31_g3(int nr, dim3 grid, hipStream_t stream, const Glynn31Args& a);
hipError_t launch_glynn31_g4(int nr, dim3 grid, hipStream_t stream, const Glynn31Args& a);
hipError_t launch_glynn31_g6(int nr, dim3 grid, hipStream_t stream, const Glynn31Args& a);

inline hipError_t launch_glynn31(int group, int nr, dim3 grid, hipStream_t stream,
                                 const Glynn31Args& a) {
    switch (group) {
        case 0: return launch_glynn31_g0(nr, grid, stream, a);
        case 2: return launch_glynn31_g2(nr, grid, stream, a);
        case 3: return launch_glynn31_g3(nr, grid, stream, a);
        case 4: return launch_glynn31_g4(nr, grid, stream, a);
        case 6: return launch_glynn31_g6(nr, grid, stream, a);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace phip
