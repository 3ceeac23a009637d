// Live-column frontier DP for 0/1 permanents on the device: the
// reference's binary_permanent_frontier / auto fast path for sparse blocks
// (perm_binary.py:389-441), with the exact counts replaced by residues
// modulo k primes (CRT on the host).
//
// A state is the set of live columns already used, as a 64-bit mask over
// live-column slots: the host gives each column a slot from its first row
// to its last, so the masks stay 64 bits wide at any order as long as at
// most 64 columns are live during a row.  A layer maps each state to its
// successors under the next row: one per available slot of the row, kept
// only when it completes every column whose last row this is, which then
// leave the mask.  Per layer:
//
//   frontier_count_kernel   successors per state;
//   hipcub ExclusiveSum     successor offsets;
//   frontier_expand_kernel  (successor mask, source state) pairs;
//   hipcub SortPairs        pairs by successor mask;
//   frontier_head_kernel +  segment heads and their exclusive sum: the
//   hipcub ExclusiveSum     distinct successors and their output slots;
//   frontier_sum_kernel     one thread per distinct successor sums its
//                           sources' residues per prime (runs are at most
//                           the row degree long).
//
// The state count of every layer is checked against the caller's cap
// before the layer's residues are formed.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "device_context.h"
#include "permhip.h"
#include "phip_arith.h"

namespace phip {

namespace {

constexpr int kFrontierBlock = 256;
constexpr int kFrontierMaxPrimes = 64;

__global__ void frontier_count_kernel(const uint64_t* __restrict__ keys, uint32_t states, uint64_t row,
                                      uint64_t expiring, uint32_t* __restrict__ counts) {
    const uint32_t i = blockIdx.x * kFrontierBlock + threadIdx.x;
    if (i >= states) return;
    const uint64_t used = keys[i];
    uint32_t c = 0;
    for (uint64_t m = row & ~used; m; m &= m - 1) {
        const uint64_t target = used | (m & (0 - m));
        c += (target & expiring) == expiring;
    }
    counts[i] = c;
}

__global__ void frontier_expand_kernel(const uint64_t* __restrict__ keys, uint32_t states, uint64_t row,
                                       uint64_t expiring, const uint32_t* __restrict__ offsets,
                                       uint64_t* __restrict__ out_keys, uint32_t* __restrict__ out_src) {
    const uint32_t i = blockIdx.x * kFrontierBlock + threadIdx.x;
    if (i >= states) return;
    const uint64_t used = keys[i];
    uint32_t at = offsets[i];
    for (uint64_t m = row & ~used; m; m &= m - 1) {
        const uint64_t target = used | (m & (0 - m));
        if ((target & expiring) != expiring) continue;
        out_keys[at] = target & ~expiring;
        out_src[at] = i;
        ++at;
    }
}

__global__ void frontier_head_kernel(const uint64_t* __restrict__ keys, uint32_t n, uint32_t* __restrict__ heads) {
    const uint32_t e = blockIdx.x * kFrontierBlock + threadIdx.x;
    if (e >= n) return;
    heads[e] = (e == 0 || keys[e] != keys[e - 1]) ? 1u : 0u;
}

// vals: [state][k] residues in [0, p_k).
__global__ void frontier_sum_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ src, uint32_t n,
                                    const uint32_t* __restrict__ heads, const uint32_t* __restrict__ slot,
                                    const uint32_t* __restrict__ vals, int k, const uint32_t* __restrict__ primes,
                                    uint64_t* __restrict__ new_keys, uint32_t* __restrict__ new_vals) {
    const uint32_t e = blockIdx.x * kFrontierBlock + threadIdx.x;
    if (e >= n || !heads[e]) return;
    const uint32_t s = slot[e];
    const uint64_t key = keys[e];
    new_keys[s] = key;
    for (int q = 0; q < k; ++q) {
        const uint64_t p = primes[q];
        uint64_t acc = 0;
        for (uint32_t f = e; f < n && keys[f] == key; ++f) {
            acc += vals[(size_t)src[f] * k + q];
            if (acc >= p) acc -= p;
        }
        new_vals[(size_t)s * k + q] = (uint32_t)acc;
    }
}

inline unsigned blocks_for(uint64_t n) { return (unsigned)((n + kFrontierBlock - 1) / kFrontierBlock); }

}  // namespace

}  // namespace phip

using phip::DeviceContext;

extern "C" int phip_binary_frontier_residues(const uint64_t* row_masks, const uint64_t* expiring, int n,
                                             int slots, const int64_t* primes, int k, int64_t max_states,
                                             int64_t* out, int64_t* info) {
    if (n < 0 || slots < 0 || slots > 64 || k < 1 || k > phip::kFrontierMaxPrimes || max_states < 1 || !out ||
        !info || (n > 0 && (!row_masks || !expiring)) || !primes) {
        phip::set_error("invalid frontier arguments (n >= 0, 0 <= slots <= 64, 1 <= k <= 64 primes, max_states >= 1)");
        return -1;
    }
    const uint64_t slot_bits = slots == 64 ? ~UINT64_C(0) : ((UINT64_C(1) << slots) - 1);
    for (int i = 0; i < n; ++i) {
        if ((row_masks[i] | expiring[i]) & ~slot_bits || (expiring[i] & ~row_masks[i])) {
            phip::set_error("frontier rows must lie within the slots and expire only their own slots");
            return -1;
        }
    }
    for (int q = 0; q < k; ++q) {
        if (primes[q] < 3 || primes[q] >= (INT64_C(1) << 31) || !(primes[q] & 1)) {
            phip::set_error("frontier moduli must be odd and lie in [3, 2^31)");
            return -1;
        }
    }
    info[0] = info[1] = 0;
    if (n == 0) {
        for (int q = 0; q < k; ++q) out[q] = 1;
        return 0;
    }
    DeviceContext* ctx = DeviceContext::current();
    if (!ctx) return -1;
    std::lock_guard<std::mutex> lock(ctx->mutex());
    hipStream_t st = ctx->stream();

    std::vector<uint32_t> hp(k), one(k);
    for (int q = 0; q < k; ++q) {
        hp[q] = (uint32_t)primes[q];
        one[q] = 1;
    }
    // Device slots 40..49: state keys and residues (40, 41 and 48, 49 in
    // turn), counts/heads (42), offsets/slots (43), pairs and sorted pairs
    // (44, 45), primes (46), hipcub scratch (47).
    uint64_t states = 1;
    auto grow = [&](int slot, size_t bytes) { return ctx->buffer(slot, std::max<size_t>(bytes, 16)); };
    auto* d_primes = (uint32_t*)grow(46, k * 4);
    auto* h_small = (uint32_t*)ctx->host_buffer(8, 64);
    if (!d_primes || !h_small) return -1;
    PHIP_HIP_TRY(hipMemcpyAsync(d_primes, hp.data(), k * 4, hipMemcpyHostToDevice, st));
    uint64_t* keys = (uint64_t*)grow(40, 8);
    uint32_t* vals = (uint32_t*)grow(41, (size_t)k * 4);
    if (!keys || !vals) return -1;
    const uint64_t zero = 0;
    PHIP_HIP_TRY(hipMemcpyAsync(keys, &zero, 8, hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(vals, one.data(), (size_t)k * 4, hipMemcpyHostToDevice, st));

    for (int pos = 0; pos < n; ++pos) {
        const uint64_t row = row_masks[pos], exp = expiring[pos];
        // Successor pairs are at most states x row degree; counts and
        // offsets are 32-bit and hipcub takes int sizes.
        if (states * (uint64_t)__builtin_popcountll(row) >= 0x7fffffffull) {
            phip::set_error("frontier layer has too many successor pairs");
            return -1;
        }
        auto* counts = (uint32_t*)grow(42, states * 4);
        auto* offsets = (uint32_t*)grow(43, states * 4);
        if (!counts || !offsets) return -1;
        hipLaunchKernelGGL(phip::frontier_count_kernel, dim3(phip::blocks_for(states)), dim3(phip::kFrontierBlock), 0,
                           st, keys, (uint32_t)states, row, exp, counts);
        PHIP_HIP_TRY(hipGetLastError());
        size_t tmp_bytes = 0;
        PHIP_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, offsets, (int)states, st));
        void* tmp = grow(47, tmp_bytes);
        if (!tmp) return -1;
        PHIP_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, counts, offsets, (int)states, st));
        PHIP_HIP_TRY(hipMemcpyAsync(h_small, offsets + states - 1, 4, hipMemcpyDeviceToHost, st));
        PHIP_HIP_TRY(hipMemcpyAsync(h_small + 1, counts + states - 1, 4, hipMemcpyDeviceToHost, st));
        PHIP_HIP_TRY(hipStreamSynchronize(st));
        const uint64_t pairs = (uint64_t)h_small[0] + h_small[1];
        if (pairs == 0) {
            for (int q = 0; q < k; ++q) out[q] = 0;
            return 0;
        }
        auto* pk = (uint64_t*)grow(44, pairs * 8 * 2);
        auto* ps = (uint32_t*)grow(45, pairs * 4 * 2);
        if (!pk || !ps) return -1;
        uint64_t* pk_sorted = pk + pairs;
        uint32_t* ps_sorted = ps + pairs;
        hipLaunchKernelGGL(phip::frontier_expand_kernel, dim3(phip::blocks_for(states)), dim3(phip::kFrontierBlock), 0,
                           st, keys, (uint32_t)states, row, exp, offsets, pk, ps);
        PHIP_HIP_TRY(hipGetLastError());
        tmp_bytes = 0;
        PHIP_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, pk, pk_sorted, ps, ps_sorted, (int)pairs,
                                                       0, std::max(slots, 1), st));
        tmp = grow(47, tmp_bytes);
        if (!tmp) return -1;
        PHIP_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, pk, pk_sorted, ps, ps_sorted, (int)pairs, 0,
                                                       std::max(slots, 1), st));
        // Heads and slots reuse the count/offset buffers, grown to the pairs.
        auto* heads = (uint32_t*)grow(42, pairs * 4);
        auto* slots = (uint32_t*)grow(43, pairs * 4);
        if (!heads || !slots) return -1;
        hipLaunchKernelGGL(phip::frontier_head_kernel, dim3(phip::blocks_for(pairs)), dim3(phip::kFrontierBlock), 0, st,
                           pk_sorted, (uint32_t)pairs, heads);
        PHIP_HIP_TRY(hipGetLastError());
        tmp_bytes = 0;
        PHIP_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, heads, slots, (int)pairs, st));
        tmp = grow(47, tmp_bytes);
        if (!tmp) return -1;
        PHIP_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, heads, slots, (int)pairs, st));
        PHIP_HIP_TRY(hipMemcpyAsync(h_small, slots + pairs - 1, 4, hipMemcpyDeviceToHost, st));
        PHIP_HIP_TRY(hipMemcpyAsync(h_small + 1, heads + pairs - 1, 4, hipMemcpyDeviceToHost, st));
        PHIP_HIP_TRY(hipStreamSynchronize(st));
        const uint64_t next = (uint64_t)h_small[0] + h_small[1];
        if (next > (uint64_t)max_states) {
            info[0] = pos + 1;
            info[1] = max_states + 1;
            return 1;
        }
        // The next layer's states alternate between slots (40, 41) and
        // (48, 49), so the current layer stays readable while it is summed.
        auto* new_keys = (uint64_t*)ctx->buffer((pos & 1) ? 40 : 48, std::max<size_t>(next * 8, 16));
        auto* new_vals = (uint32_t*)ctx->buffer((pos & 1) ? 41 : 49, std::max<size_t>(next * k * 4, 16));
        if (!new_keys || !new_vals) return -1;
        hipLaunchKernelGGL(phip::frontier_sum_kernel, dim3(phip::blocks_for(pairs)), dim3(phip::kFrontierBlock), 0, st,
                           pk_sorted, ps_sorted, (uint32_t)pairs, heads, slots, vals, k, d_primes, new_keys, new_vals);
        PHIP_HIP_TRY(hipGetLastError());
        keys = new_keys;
        vals = new_vals;
        states = next;
    }
    // Every column has expired after the last row: the only state is 0.
    std::vector<uint32_t> res((size_t)states * k);
    std::vector<uint64_t> fk(states);
    PHIP_HIP_TRY(hipMemcpyAsync(fk.data(), keys, states * 8, hipMemcpyDeviceToHost, st));
    PHIP_HIP_TRY(hipMemcpyAsync(res.data(), vals, (size_t)states * k * 4, hipMemcpyDeviceToHost, st));
    PHIP_HIP_TRY(hipStreamSynchronize(st));
    for (int q = 0; q < k; ++q) out[q] = 0;
    for (uint64_t i = 0; i < states; ++i) {
        if (fk[i] == 0) {
            for (int q = 0; q < k; ++q) out[q] = res[(size_t)i * k + q];
        }
    }
    return 0;
}
