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
// Small layers take a faster path: frontier_small_kernel, one workgroup of
// 1024 threads, runs consecutive layers with the current layer, a hash
// table of successor keys and the successor sums all in LDS, and returns
// when a layer would not fit (more than kSmallStates states or kSmallPairs
// successor pairs, or more than kSmallPrimes moduli), when a layer passes
// the cap, or at the end.  The host then continues with the kernels above
// for one layer, and hands back to the small kernel when the layers fit
// again.  A layer costs a few barriers instead of eight launches and two
// host synchronisations.
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

constexpr int kSmallThreads = 1024;
constexpr int kSmallStates = 1024;  // states of a layer held in LDS
constexpr int kSmallPairs = 4096;   // successor pairs of one layer
constexpr int kSmallHash = 8192;    // hash slots (load at most one half)
constexpr int kSmallPrimes = 4;
constexpr uint64_t kEmptyKey = ~UINT64_C(0);  // no state has every slot bit set with 64 slots in use
constexpr uint32_t kSmallDone = 0;   // every row processed
constexpr uint32_t kSmallCap = 1;    // the layer of row `pos` passes the cap
constexpr uint32_t kSmallLarge = 2;  // the layer of row `pos` does not fit; current layer returned
constexpr uint32_t kSmallEmpty = 3;  // no successor: the count is zero

// Exclusive prefix sum over the workgroup; *total receives the sum.
__device__ __forceinline__ uint32_t small_block_scan(uint32_t v, uint32_t* wave_sums, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wave_sums[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kSmallThreads / 64; ++w) {
        const uint32_t t = wave_sums[w];
        if (w < wave) before += t;
        all += t;
    }
    *total = all;
    __syncthreads();
    return before + x - v;
}

// Layers pos0, pos0 + 1, ... of the DP in one workgroup.  state[0] holds
// the current layer's size on entry (at most kSmallStates, with its keys and
// residues in keys / vals, k per state); on exit state[0..2] = (size of the
// returned layer, first unprocessed row, status) and keys / vals hold that
// layer.  k <= kSmallPrimes.
__global__ __launch_bounds__(kSmallThreads) void frontier_small_kernel(
    const uint64_t* __restrict__ rows, const uint64_t* __restrict__ expiring, int n, int pos0, int k,
    const uint32_t* __restrict__ primes, uint64_t max_states, uint64_t* __restrict__ keys,
    uint32_t* __restrict__ vals, uint32_t* __restrict__ state) {
    __shared__ uint64_t sk[kSmallStates];
    __shared__ uint32_t sv[kSmallStates * kSmallPrimes];
    __shared__ unsigned long long hk[kSmallHash];
    __shared__ unsigned long long acc[kSmallStates * kSmallPrimes];
    __shared__ uint16_t psrc[kSmallPairs];
    __shared__ uint16_t pslot[kSmallPairs];
    __shared__ uint32_t wave_sums[kSmallThreads / 64];
    __shared__ uint32_t npairs, overflow;
    const uint32_t tid = threadIdx.x;
    uint32_t count = state[0];
    for (uint32_t i = tid; i < count; i += kSmallThreads) {
        sk[i] = keys[i];
        for (int q = 0; q < k; ++q) sv[i * kSmallPrimes + q] = vals[(size_t)i * k + q];
    }
    uint32_t p[kSmallPrimes];
#pragma unroll
    for (int q = 0; q < kSmallPrimes; ++q) p[q] = q < k ? primes[q] : 1u;
    __syncthreads();
    int pos = pos0;
    uint32_t status = kSmallDone;
    for (; pos < n; ++pos) {
        const uint64_t row = rows[pos], exp = expiring[pos];
        if ((uint64_t)count * (uint64_t)__popcll(row) > (uint64_t)kSmallPairs) {
            status = kSmallLarge;
            break;
        }
        for (int s = tid; s < kSmallHash; s += kSmallThreads) hk[s] = kEmptyKey;
        for (int s = tid; s < kSmallStates * kSmallPrimes; s += kSmallThreads) acc[s] = 0;
        if (tid == 0) npairs = overflow = 0;
        __syncthreads();
        // Successor pairs: insert each successor key (linear probing) and
        // record (source, slot).
        for (uint32_t i = tid; i < count; i += kSmallThreads) {
            const uint64_t used = sk[i];
            for (uint64_t m = row & ~used; m; m &= m - 1) {
                const uint64_t target = used | (m & (0 - m));
                if ((target & exp) != exp) continue;
                const unsigned long long key = target & ~exp;
                uint32_t h = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 51);  // 13 bits
                int probe = 0;
                for (; probe < kSmallHash; ++probe) {
                    const unsigned long long old = atomicCAS(&hk[h], kEmptyKey, key);
                    if (old == kEmptyKey || old == key) break;
                    h = (h + 1) & (kSmallHash - 1);
                }
                if (probe == kSmallHash) overflow = 1;  // not reachable: at most kSmallPairs keys
                const uint32_t at = atomicAdd(&npairs, 1u);
                psrc[at] = (uint16_t)i;
                pslot[at] = (uint16_t)h;
            }
        }
        __syncthreads();
        // Dense numbering of the occupied slots (8 consecutive per thread).
        constexpr int kPer = kSmallHash / kSmallThreads;
        uint32_t local = 0;
#pragma unroll
        for (int j = 0; j < kPer; ++j) local += hk[tid * kPer + j] != kEmptyKey;
        uint32_t distinct = 0;
        const uint32_t base = small_block_scan(local, wave_sums, &distinct);
        if (distinct > max_states) {
            status = kSmallCap;
            break;
        }
        if (distinct == 0) {
            status = kSmallEmpty;
            count = 0;
            break;
        }
        if (distinct > (uint32_t)kSmallStates || overflow) {
            status = kSmallLarge;
            break;
        }
        // The current keys are no longer read: the successors take their
        // places, and each occupied slot now holds its dense index.
        uint32_t d = base;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int s = tid * kPer + j;
            const unsigned long long key = hk[s];
            if (key != kEmptyKey) {
                sk[d] = key;
                hk[s] = d;
                ++d;
            }
        }
        __syncthreads();
        const uint32_t np = npairs;
        for (uint32_t a = tid; a < np; a += kSmallThreads) {
            const uint32_t dst = (uint32_t)hk[pslot[a]];
            const uint32_t src = psrc[a];
            for (int q = 0; q < k; ++q)
                atomicAdd(&acc[dst * kSmallPrimes + q], (unsigned long long)sv[src * kSmallPrimes + q]);
        }
        __syncthreads();
        for (uint32_t e = tid; e < distinct; e += kSmallThreads) {
            for (int q = 0; q < k; ++q)
                sv[e * kSmallPrimes + q] = (uint32_t)(acc[e * kSmallPrimes + q] % p[q]);
        }
        count = distinct;
        __syncthreads();
    }
    for (uint32_t i = tid; i < count; i += kSmallThreads) {
        keys[i] = sk[i];
        for (int q = 0; q < k; ++q) vals[(size_t)i * k + q] = sv[i * kSmallPrimes + q];
    }
    if (tid == 0) {
        state[0] = count;
        state[1] = (uint32_t)pos;
        state[2] = status;
    }
}

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
    // Device slots 40..52: state keys and residues (40, 41 and 48, 49 in
    // turn), counts/heads (42), offsets/output slots (43), pairs and sorted
    // pairs (44, 45), primes (46), hipcub scratch (47), rows and expiring
    // masks (50, 51), small-kernel state (52).
    uint64_t states = 1;
    auto grow = [&](int slot, size_t bytes) { return ctx->buffer(slot, std::max<size_t>(bytes, 16)); };
    auto* d_primes = (uint32_t*)grow(46, k * 4);
    auto* h_small = (uint32_t*)ctx->host_buffer(8, 64);
    auto* d_rows = (uint64_t*)grow(50, (size_t)n * 8);
    auto* d_exp = (uint64_t*)grow(51, (size_t)n * 8);
    auto* d_state = (uint32_t*)grow(52, 16);
    if (!d_primes || !h_small || !d_rows || !d_exp || !d_state) return -1;
    PHIP_HIP_TRY(hipMemcpyAsync(d_primes, hp.data(), k * 4, hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d_rows, row_masks, (size_t)n * 8, hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d_exp, expiring, (size_t)n * 8, hipMemcpyHostToDevice, st));
    // Both layer slots hold at least one small-kernel layer, so the small
    // kernel, which works in place, never needs its slot grown.
    const size_t small_keys = (size_t)phip::kSmallStates * 8, small_vals = (size_t)phip::kSmallStates * k * 4;
    if (!grow(48, small_keys) || !grow(49, small_vals)) return -1;
    int cur = 40;  // slot of the current layer's keys; its residues are in cur + 1
    uint64_t* keys = (uint64_t*)grow(40, small_keys);
    uint32_t* vals = (uint32_t*)grow(41, small_vals);
    if (!keys || !vals) return -1;
    const uint64_t zero = 0;
    PHIP_HIP_TRY(hipMemcpyAsync(keys, &zero, 8, hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(vals, one.data(), (size_t)k * 4, hipMemcpyHostToDevice, st));

    // The small kernel needs k <= kSmallPrimes and fewer than 64 slots (an
    // all-ones key marks an empty hash slot).
    const bool small_ok = k <= phip::kSmallPrimes && slots < 64;
    bool force_large = false;
    int pos = 0;
    while (pos < n) {
        const uint64_t row = row_masks[pos], exp = expiring[pos];
        const uint64_t degree = (uint64_t)__builtin_popcountll(row);
        if (small_ok && !force_large && states <= (uint64_t)phip::kSmallStates &&
            states * degree <= (uint64_t)phip::kSmallPairs) {
            h_small[0] = (uint32_t)states;
            PHIP_HIP_TRY(hipMemcpyAsync(d_state, h_small, 4, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(phip::frontier_small_kernel, dim3(1), dim3(phip::kSmallThreads), 0, st, d_rows, d_exp,
                               n, pos, k, d_primes, (uint64_t)max_states, keys, vals, d_state);
            PHIP_HIP_TRY(hipGetLastError());
            PHIP_HIP_TRY(hipMemcpyAsync(h_small, d_state, 12, hipMemcpyDeviceToHost, st));
            PHIP_HIP_TRY(hipStreamSynchronize(st));
            const uint32_t reached = h_small[1], status = h_small[2];
            states = h_small[0];
            if (status == phip::kSmallCap) {
                info[0] = (int64_t)reached + 1;
                info[1] = max_states + 1;
                return 1;
            }
            if (status == phip::kSmallEmpty) {
                for (int q = 0; q < k; ++q) out[q] = 0;
                return 0;
            }
            // A layer that does not fit goes through the kernels below once.
            force_large = status == phip::kSmallLarge;
            pos = (int)reached;
            continue;
        }
        force_large = false;
        // Successor pairs are at most states x row degree; counts and
        // offsets are 32-bit and hipcub takes int sizes.
        if (states * degree >= 0x7fffffffull) {
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
        // Heads and output slots reuse the count/offset buffers, grown to
        // the pairs.
        auto* heads = (uint32_t*)grow(42, pairs * 4);
        auto* out_slot = (uint32_t*)grow(43, pairs * 4);
        if (!heads || !out_slot) return -1;
        hipLaunchKernelGGL(phip::frontier_head_kernel, dim3(phip::blocks_for(pairs)), dim3(phip::kFrontierBlock), 0, st,
                           pk_sorted, (uint32_t)pairs, heads);
        PHIP_HIP_TRY(hipGetLastError());
        tmp_bytes = 0;
        PHIP_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, heads, out_slot, (int)pairs, st));
        tmp = grow(47, tmp_bytes);
        if (!tmp) return -1;
        PHIP_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, heads, out_slot, (int)pairs, st));
        PHIP_HIP_TRY(hipMemcpyAsync(h_small, out_slot + pairs - 1, 4, hipMemcpyDeviceToHost, st));
        PHIP_HIP_TRY(hipMemcpyAsync(h_small + 1, heads + pairs - 1, 4, hipMemcpyDeviceToHost, st));
        PHIP_HIP_TRY(hipStreamSynchronize(st));
        const uint64_t next = (uint64_t)h_small[0] + h_small[1];
        if (next > (uint64_t)max_states) {
            info[0] = pos + 1;
            info[1] = max_states + 1;
            return 1;
        }
        // The next layer goes to the other pair of slots, so the current
        // layer stays readable while it is summed.
        const int other = cur == 40 ? 48 : 40;
        auto* new_keys = (uint64_t*)ctx->buffer(other, std::max<size_t>(next * 8, small_keys));
        auto* new_vals = (uint32_t*)ctx->buffer(other + 1, std::max<size_t>(next * k * 4, small_vals));
        if (!new_keys || !new_vals) return -1;
        hipLaunchKernelGGL(phip::frontier_sum_kernel, dim3(phip::blocks_for(pairs)), dim3(phip::kFrontierBlock), 0, st,
                           pk_sorted, ps_sorted, (uint32_t)pairs, heads, out_slot, vals, k, d_primes, new_keys,
                           new_vals);
        PHIP_HIP_TRY(hipGetLastError());
        cur = other;
        keys = new_keys;
        vals = new_vals;
        states = next;
        ++pos;
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
