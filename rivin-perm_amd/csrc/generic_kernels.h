// Kernels for the even moduli the Glynn kernels cannot serve, so that the
// modular entry point accepts every modulus the reference's perm_c_ext does
// (any 0 < m < 2^63, perm_c_omp.c:876-880); odd moduli of any width run on
// the Glynn kernels of batch.hip and wide.hip:
//
//   ryser_u64_kernel Ryser's formula in wrapping 64-bit arithmetic, i.e. the
//                    permanent modulo 2^64 (perm_c_omp.c:666-770 restated
//                    over Z/2^64).  An even modulus 2^e * q is served by this
//                    kernel (mask to e bits) together with an odd-q residue,
//                    combined by CRT on the host.
//   sum_mod_kernel   one workgroup per modulus sums per-lane or per-block
//                    partials modulo that modulus (tree in LDS, no atomics).
//
// These paths exist for completeness of the modular contract (even moduli);
// the hot path is the persistent kernel of batch.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "phip_arith.h"

namespace phip {

constexpr int kGenericBlock = 256;
constexpr int kMaxOrder = 63;

// A[n][n] column-major, entries as two's-complement 64-bit words.  Each lane
// covers 2^cbits Ryser subsets; partials[lane] = sum of (-1)^|S| prod rs
// over its subsets, modulo 2^64.
__global__ __launch_bounds__(kGenericBlock) void ryser_u64_kernel(
    const uint64_t* __restrict__ A, int n, int cbits, uint64_t lanes,
    uint64_t* __restrict__ partials) {
    const uint64_t lane = (uint64_t)blockIdx.x * kGenericBlock + threadIdx.x;
    if (lane >= lanes) return;
    const uint64_t base = lane << cbits;
    uint64_t gray = base ^ (base >> 1);
    uint64_t rs[kMaxOrder];
    for (int i = 0; i < n; ++i) rs[i] = 0;
    for (int j = 0; j < n; ++j) {
        if ((gray >> j) & 1u) {
            for (int i = 0; i < n; ++i) rs[i] += A[j * n + i];
        }
    }
    bool negative = __popcll(gray) & 1;
    uint64_t acc = 0;
    const uint64_t states = 1ull << cbits;
    for (uint64_t t = 0; t < states; ++t) {
        if (t != 0) {
            const int j = __builtin_ctzll(t);
            const uint64_t s = base + t;
            gray = s ^ (s >> 1);
            const bool add = (gray >> j) & 1u;
            for (int i = 0; i < n; ++i) rs[i] = add ? rs[i] + A[j * n + i] : rs[i] - A[j * n + i];
            negative = !negative;
        }
        uint64_t prod = rs[0];
        for (int i = 1; i < n; ++i) prod *= rs[i];
        acc = negative ? acc - prod : acc + prod;
    }
    partials[lane] = acc;
}

// out[q] = sum_{i < count} partials[q][i] mod moduli[q]; a modulus of 0
// means 2^64 (plain wrapping sum).
__global__ __launch_bounds__(256) void sum_mod_kernel(const uint64_t* __restrict__ partials,
                                                      uint64_t count,
                                                      const uint64_t* __restrict__ moduli,
                                                      uint64_t* __restrict__ out) {
    __shared__ uint64_t tree[256];
    const int q = blockIdx.x;
    const uint64_t p = moduli[q];
    const uint64_t* __restrict__ src = partials + (size_t)q * count;
    uint64_t s = 0;
    for (uint64_t i = threadIdx.x; i < count; i += 256) {
        if (p == 0) {
            s += src[i];
        } else {
            s = add_mod_u64(s, src[i] % p, p);
        }
    }
    tree[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const uint64_t a = tree[threadIdx.x];
            const uint64_t b = tree[threadIdx.x + w];
            tree[threadIdx.x] = (p == 0) ? a + b : add_mod_u64(a, b, p);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = tree[0];
}

}  // namespace phip
