// Residue engine: splits each modulus m = 2^e * q into an odd part served by
// the Glynn kernels and a power-of-two part served by the wrapping Ryser
// kernel, runs every odd part of one arithmetic class together, and
// reassembles the residues (CRT for even moduli).
//
// Host preparation follows the reference's normalisation of entries before
// fixed-width arithmetic (perm_c_omp.c:912-916, normalize_mod_i64).
#include "residues.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <string>

#include "batch.h"
#include "generic_kernels.h"
#include "launchers.h"
#include "phip_arith.h"
#include "tables.h"
#include "wide.h"

namespace phip {

Tuning& tuning() {
    static Tuning t;
    return t;
}

namespace {

int bit_length(uint64_t v) { return v ? 64 - __builtin_clzll(v) : 0; }

int ceil_log2(uint64_t v) {
    int l = 0;
    while ((1ull << l) < v) ++l;
    return l;
}

uint64_t pow2_mod(int e, uint64_t q) { return host_powmod(2, (uint64_t)e, q); }

const int64_t* matrix_of(const int64_t* A, int64_t stride, int index) { return A + (int64_t)index * stride; }

// ---------------------------------------------------------------------
// Ryser modulo 2^64.
// ---------------------------------------------------------------------
int run_ryser_u64(DeviceContext* ctx, const int64_t* M, int n, uint64_t* result) {
    std::vector<uint64_t> cols((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) cols[(size_t)j * n + i] = (uint64_t)M[(size_t)i * n + j];
    const int cbits = std::max(0, n - 14);
    const uint64_t lanes = 1ull << (n - cbits);
    const uint64_t gx = (lanes + kGenericBlock - 1) / kGenericBlock;
    hipStream_t st = ctx->stream();
    auto* d_a = (uint64_t*)ctx->buffer(13, cols.size() * 8);
    auto* d_part = (uint64_t*)ctx->buffer(14, lanes * 8);
    auto* d_sum = (uint64_t*)ctx->buffer(15, 16);
    auto* h_sum = (uint64_t*)ctx->host_buffer(2, 16);
    if (!d_a || !d_part || !d_sum || !h_sum) return -1;
    const uint64_t zero = 0;
    PHIP_HIP_TRY(hipMemcpyAsync(d_a, cols.data(), cols.size() * 8, hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d_sum + 1, &zero, 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(ryser_u64_kernel, dim3((unsigned)gx), dim3(kGenericBlock), 0, st, d_a, n, cbits,
                       lanes, d_part);
    PHIP_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(sum_mod_kernel, dim3(1), dim3(256), 0, st, d_part, lanes, d_sum + 1, d_sum);
    PHIP_HIP_TRY(hipGetLastError());
    PHIP_HIP_TRY(hipMemcpyAsync(h_sum, d_sum, 8, hipMemcpyDeviceToHost, st));
    PHIP_HIP_TRY(hipStreamSynchronize(st));
    // Ryser's (-1)^n factor (perm_c_omp.c:767-770).
    *result = (n & 1) ? (uint64_t)0 - h_sum[0] : h_sum[0];
    return 0;
}

}  // namespace

bool is_prime_u64(uint64_t v) {
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (v < 2) return false;
    for (uint64_t p : small) {
        if (v == p) return true;
        if (v % p == 0) return false;
    }
    uint64_t d = v - 1;
    int r = 0;
    while (!(d & 1)) {
        d >>= 1;
        ++r;
    }
    // Bases proven sufficient for every 64-bit integer (Sinclair's set).
    static const uint64_t bases[] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    for (uint64_t a : bases) {
        a %= v;
        if (a == 0) continue;
        uint64_t x = host_powmod(a, d, v);
        if (x == 1 || x == v - 1) continue;
        bool composite = true;
        for (int i = 1; i < r; ++i) {
            x = host_mulmod(x, x, v);
            if (x == v - 1) {
                composite = false;
                break;
            }
        }
        if (composite) return false;
    }
    return true;
}

int compute_residues(DeviceContext* ctx, const int64_t* A, int64_t stride, int n, const int64_t* moduli,
                     int k, int64_t* out) {
    if (n < 0 || n >= 64) {
        set_error("matrix dimension must satisfy 0 <= n < 64");
        return -1;
    }
    if (k < 0 || (k > 0 && (!moduli || !out))) {
        set_error("invalid modulus list");
        return -1;
    }
    if (n > 0 && !A) {
        set_error("matrix pointer is null");
        return -1;
    }
    for (int i = 0; i < k; ++i) {
        if (moduli[i] <= 0) {
            set_error("moduli must be positive signed 64-bit integers");
            return -1;
        }
    }
    if (k == 0) return 0;
    if (n == 0) {
        for (int i = 0; i < k; ++i) out[i] = 1 % moduli[i];
        return 0;
    }
    if (n == 1) {
        for (int i = 0; i < k; ++i) out[i] = host_floor_mod(matrix_of(A, stride, i)[0], moduli[i]);
        return 0;
    }

    std::vector<uint64_t> odd_residue(k, 0);
    // Odd parts below 2^31 at orders [5, 46] run on the persistent Glynn
    // engine as a one-matrix plan; moduli whose tables coincide share one
    // Gray walk.  The rest run on the wide Glynn kernels (32-bit Montgomery,
    // 61-bit pseudo-Mersenne or 64-bit Montgomery by modulus).
    std::vector<PlanInput> plan_in;
    std::vector<int> plan_index;
    std::vector<WideJob> wide;
    for (int i = 0; i < k; ++i) {
        const uint64_t m = (uint64_t)moduli[i];
        const int e = __builtin_ctzll(m);
        const uint64_t q = m >> e;
        if (q <= 1) continue;
        if (q < (1ull << 31) && n >= kGlynnMinOrder && n <= kGlynnMaxOrder) {
            plan_in.push_back({matrix_of(A, stride, i), n, (uint32_t)q, 0});
            plan_index.push_back(i);
        } else {
            wide.push_back({i, q});
        }
    }
    if (!plan_in.empty()) {
        std::vector<int64_t> res(plan_in.size());
        if (run_plan_in_context(ctx, plan_in, res.data()) != 0) return -1;
        for (size_t j = 0; j < plan_in.size(); ++j) odd_residue[plan_index[j]] = (uint64_t)res[j];
    }
    if (!wide.empty() && run_wide_glynn(ctx, A, stride, n, wide, odd_residue) != 0) return -1;

    bool have_two = false;
    uint64_t two_residue = 0;
    for (int i = 0; i < k; ++i) {
        const uint64_t m = (uint64_t)moduli[i];
        const int e = __builtin_ctzll(m);
        const uint64_t q = m >> e;
        if (e == 0) {
            out[i] = (int64_t)odd_residue[i];
            continue;
        }
        uint64_t r2;
        if (stride == 0) {
            if (!have_two) {
                if (run_ryser_u64(ctx, A, n, &two_residue) != 0) return -1;
                have_two = true;
            }
            r2 = two_residue;
        } else {
            if (run_ryser_u64(ctx, matrix_of(A, stride, i), n, &r2) != 0) return -1;
        }
        const uint64_t pow2 = 1ull << e;  // e <= 62
        r2 &= pow2 - 1;
        if (q == 1) {
            out[i] = (int64_t)r2;
            continue;
        }
        // x = r2 + 2^e * t with t = (r_q - r2) * 2^-e mod q.
        const uint64_t rq = odd_residue[i];
        const uint64_t diff = (rq + q - (r2 % q)) % q;
        const uint64_t inv = host_powmod((q + 1) / 2, (uint64_t)e, q);
        const uint64_t t = host_mulmod(diff, inv, q);
        out[i] = (int64_t)(r2 + pow2 * t);
    }
    return 0;
}

}  // namespace phip
