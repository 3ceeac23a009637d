// extern "C" entry points of libpermhip.so (declared in include/permhip.h).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "device_context.h"
#include "permhip.h"
#include "phip_arith.h"
#include "necklace.h"
#include "affine.h"
#include "residues.h"
#include "batch.h"

namespace phip {
int mod3_from_columns(DeviceContext* ctx, const uint64_t* mags, const uint64_t* sgns, int n);
}

using phip::DeviceContext;

namespace {

// Validation shared by the binary entry points, mirroring the reference's
// native checks (perm_binary_omp.c:422-452, perm_binary_cuda.cu:314-340).
bool validate_binary(const uint8_t* A, int n, const int64_t* primes, int k) {
    if (n < 0 || n > 63) {
        phip::set_error("native binary backend supports order at most 63");
        return false;
    }
    if (n > 0 && !A) {
        phip::set_error("matrix pointer is null");
        return false;
    }
    if (k < 0 || (k > 0 && !primes)) {
        phip::set_error("invalid prime list");
        return false;
    }
    for (int i = 0; i < k; ++i) {
        const int64_t p = primes[i];
        if (p <= 1) {
            phip::set_error("p must be greater than 1");
            return false;
        }
        if (p == 2) {
            phip::set_error("p must be odd");
            return false;
        }
        if (p <= n) {
            phip::set_error("p must be greater than the matrix order");
            return false;
        }
        if (!phip::is_prime_u64((uint64_t)p)) {
            phip::set_error("p must be prime");
            return false;
        }
    }
    for (int i = 0; i < n * n; ++i) {
        if (A[i] > 1) {
            phip::set_error("matrix entries must be 0 or 1");
            return false;
        }
    }
    return true;
}

}  // namespace

extern "C" {

int phip_available(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) return 0;
    return count > 0 ? 1 : 0;
}

int phip_last_error(char* buf, size_t len) {
    const std::string& e = phip::last_error();
    if (buf && len > 0) {
        const size_t m = e.size() < len - 1 ? e.size() : len - 1;
        memcpy(buf, e.data(), m);
        buf[m] = '\0';
    }
    return (int)e.size();
}

int phip_device_info(char* buf, size_t len) {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        phip::set_error("no HIP device available");
        return -1;
    }
    std::string s = std::string(prop.name) + " " + prop.gcnArchName + " CUs=" +
                    std::to_string(prop.multiProcessorCount) + " clock_kHz=" + std::to_string(prop.clockRate);
    if (buf && len > 0) {
        const size_t m = s.size() < len - 1 ? s.size() : len - 1;
        memcpy(buf, s.data(), m);
        buf[m] = '\0';
    }
    return 0;
}

int phip_set_tuning(int key, int64_t value) {
    phip::Tuning& t = phip::tuning();
    switch (key) {
        case PHIP_TUNE_LANES_LOG2:
            if (value < 8 || value > 26) break;
            t.lanes_log2 = (int)value;
            return 0;
        case PHIP_TUNE_MIN_CBITS:
            if (value < 4 || value > 30) break;
            t.min_cbits = (int)value;
            return 0;
        case PHIP_TUNE_BATCH_ITEMS_LOG2:
            if (value < 6 || value > 24) break;
            t.batch_items_log2 = (int)value;
            return 0;
        case PHIP_TUNE_BATCH_BLOCKS:
            if (value < 1 || value > 64) break;
            t.batch_blocks_per_cu = (int)value;
            return 0;
        case PHIP_TUNE_NECKLACE_MIN_ORDER:
            if (value < 5 || value > 64) break;
            t.necklace_min_order = (int)value;
            return 0;
        case PHIP_TUNE_NECKLACE_DEPTH:
            if (value < 2 || value > 30) break;
            t.necklace_depth = (int)value;
            return 0;
        case PHIP_TUNE_WIDE_MIN_TIER:
            if (value < 8 || value > 64 || value % 4 != 0) break;
            t.wide_min_tier = (int)value;
            return 0;
        case PHIP_TUNE_AFFINE_MIN_TIER:
            if (value < 8 || value > 64 || value % 4 != 0 || (value > 40 && value % 8 != 0)) break;
            t.affine_min_tier = (int)value;
            return 0;
        case PHIP_TUNE_SKIP_ZEROS_PERMILLE:
            if (value < 0 || value > 1000) break;
            t.skip_zeros_permille = (int)value;
            return 0;
        case PHIP_TUNE_EXACT_MIN_PRIMES:
            if (value < 0 || value > 1024) break;
            t.exact_min_primes = (int)value;
            return 0;
        case PHIP_TUNE_EXACT_WIDE_MIN_PRIMES:
            if (value < 0 || value > 1024) break;
            t.exact_wide_min_primes = (int)value;
            return 0;
        case PHIP_TUNE_EXACT_COMPACT:
            if (value < 0 || value > 1) break;
            t.exact_compact = (int)value;
            return 0;
        case PHIP_TUNE_SKIP_BLOCKS:
            if (value < 0 || value > 1) break;
            t.skip_blocks = (int)value;
            return 0;
        case PHIP_TUNE_MOD3_BLOCKS:
            if (value < 0 || value > 1) break;
            t.mod3_blocks = (int)value;
            return 0;
        case PHIP_TUNE_EXACT_SIGNED:
            if (value < 0 || value > 1) break;
            t.exact_signed = (int)value;
            return 0;
        case PHIP_TUNE_SKIP_EXACT_MIN_PRIMES:
            if (value < 0 || value > 1024) break;
            t.skip_exact_min_primes = (int)value;
            return 0;
        default:
            phip::set_error("unknown tuning key");
            return -1;
    }
    phip::set_error("tuning value out of range");
    return -1;
}

int64_t phip_get_tuning(int key) {
    const phip::Tuning& t = phip::tuning();
    switch (key) {
        case PHIP_TUNE_LANES_LOG2: return t.lanes_log2;
        case PHIP_TUNE_MIN_CBITS: return t.min_cbits;
        case PHIP_TUNE_BATCH_ITEMS_LOG2: return t.batch_items_log2;
        case PHIP_TUNE_BATCH_BLOCKS: return t.batch_blocks_per_cu;
        case PHIP_TUNE_NECKLACE_MIN_ORDER: return t.necklace_min_order;
        case PHIP_TUNE_NECKLACE_DEPTH: return t.necklace_depth;
        case PHIP_TUNE_WIDE_MIN_TIER: return t.wide_min_tier;
        case PHIP_TUNE_AFFINE_MIN_TIER: return t.affine_min_tier;
        case PHIP_TUNE_SKIP_ZEROS_PERMILLE: return t.skip_zeros_permille;
        case PHIP_TUNE_EXACT_MIN_PRIMES: return t.exact_min_primes;
        case PHIP_TUNE_EXACT_WIDE_MIN_PRIMES: return t.exact_wide_min_primes;
        case PHIP_TUNE_EXACT_COMPACT: return t.exact_compact;
        case PHIP_TUNE_SKIP_BLOCKS: return t.skip_blocks;
        case PHIP_TUNE_MOD3_BLOCKS: return t.mod3_blocks;
        case PHIP_TUNE_SKIP_EXACT_MIN_PRIMES: return t.skip_exact_min_primes;
        case PHIP_TUNE_EXACT_SIGNED: return t.exact_signed;
        default:
            phip::set_error("unknown tuning key");
            return -1;
    }
}

int phip_profile_enable(int on) {
    DeviceContext* ctx = DeviceContext::current();
    if (!ctx) return -1;
    std::lock_guard<std::mutex> lock(ctx->mutex());
    if (on && !ctx->prof_start) {
        PHIP_HIP_TRY(hipEventCreate(&ctx->prof_start));
        PHIP_HIP_TRY(hipEventCreate(&ctx->prof_stop));
    }
    ctx->profiling = on != 0;
    ctx->profiled_ms = 0;
    ctx->profiled_launches = 0;
    return 0;
}

int phip_profile_read(double* kernel_ms, uint64_t* launches) {
    DeviceContext* ctx = DeviceContext::current();
    if (!ctx) return -1;
    std::lock_guard<std::mutex> lock(ctx->mutex());
    if (kernel_ms) *kernel_ms = ctx->profiled_ms;
    if (launches) *launches = ctx->profiled_launches;
    return 0;
}

int phip_permanent_residues(const int64_t* A, int64_t prime_stride, int n, const int64_t* moduli, int k,
                            int64_t* out) {
    DeviceContext* ctx = DeviceContext::current();
    if (!ctx) return -1;
    std::lock_guard<std::mutex> lock(ctx->mutex());
    return phip::compute_residues(ctx, A, prime_stride, n, moduli, k, out);
}

int phip_affine_permanent_residues(const int64_t* C, const int64_t* L, int64_t prime_stride, int n,
                                   const int64_t* moduli, int k, int64_t* out) {
    DeviceContext* ctx = DeviceContext::current();
    if (!ctx) return -1;
    std::lock_guard<std::mutex> lock(ctx->mutex());
    return phip::affine_permanent_residues(ctx, C, L, prime_stride, n, moduli, k, phip::tuning().affine_min_tier,
                                           out);
}

int64_t phip_permanent_modp(const int64_t* A, int n, int64_t p) {
    int64_t r = -1;
    if (phip_permanent_residues(A, 0, n, &p, 1, &r) != 0) return -1;
    return r;
}

int64_t phip_schur_permanent_modp(int n, int64_t omega, int64_t p) {
    if (n < 0 || n >= 64) {
        phip::set_error("n must satisfy 0 <= n < 64");
        return -1;
    }
    if (p <= 0) {
        phip::set_error("p must be a positive signed 64-bit integer");
        return -1;
    }
    if (n <= 1) return 1 % p;
    if (p == 1) return 0;
    const uint64_t P = (uint64_t)p;
    const uint64_t w = (uint64_t)phip::host_floor_mod(omega, p);
    // Cofactor identity for odd n when omega^n == 1 (perm_c_omp.c:964-1025):
    // every permanental cofactor of the all-ones first row is equal.
    const bool cofactor = n >= 3 && (n & 1) && (p & 1) && phip::host_powmod(w, (uint64_t)n, P) == 1 % P;
    if (cofactor && n >= phip::tuning().necklace_min_order && phip::necklace_supported(n, P) &&
        phip::is_prime_u64(P)) {
        DeviceContext* ctx = DeviceContext::current();
        if (!ctx) return -1;
        std::lock_guard<std::mutex> lock(ctx->mutex());
        int64_t minor = -1;
        if (phip::necklace_minor_modp(ctx, n, w, P, &minor) != 0) return -1;
        return (int64_t)phip::host_mulmod((uint64_t)minor, (uint64_t)n % P, P);
    }
    const int m = cofactor ? n - 1 : n;
    const uint64_t first = cofactor ? w : 1 % P;  // omega^(row index) of the first row used
    std::vector<int64_t> M((size_t)m * m);
    uint64_t wj = first;
    for (int r = 0; r < m; ++r) {
        uint64_t pw = cofactor ? wj : 1 % P;
        for (int c = 0; c < m; ++c) {
            M[(size_t)r * m + c] = (int64_t)pw;
            pw = phip::host_mulmod(pw, wj, P);
        }
        wj = phip::host_mulmod(wj, w, P);
    }
    int64_t r = -1;
    if (phip_permanent_residues(M.data(), 0, m, &p, 1, &r) != 0) return -1;
    if (cofactor) r = (int64_t)phip::host_mulmod((uint64_t)r, (uint64_t)n % P, P);
    return r;
}

int phip_schur_necklace_orbit_stats(int n, uint64_t* representatives, uint64_t* words) {
    if (!representatives || !words) {
        phip::set_error("null argument");
        return -1;
    }
    if (n < 3 || n >= 64 || !(n & 1) || !phip::is_prime_u64((uint64_t)n)) {
        phip::set_error("n must be an odd prime satisfying 3 <= n < 64");
        return -1;
    }
    DeviceContext* ctx = DeviceContext::current();
    if (!ctx) return -1;
    std::lock_guard<std::mutex> lock(ctx->mutex());
    return phip::necklace_orbit_stats(ctx, n, representatives, words);
}

int phip_binary_permanent_residues(const uint8_t* A, int n, const int64_t* primes, int k, int64_t* out) {
    if (!validate_binary(A, n, primes, k)) return -1;
    if (k > 0 && !out) {
        phip::set_error("output pointer is null");
        return -1;
    }
    std::vector<int64_t> M((size_t)n * n);
    for (int i = 0; i < n * n; ++i) M[i] = A[i];
    return phip_permanent_residues(n ? M.data() : nullptr, 0, n, primes, k, out);
}

int64_t phip_binary_permanent_modp(const uint8_t* A, int n, int64_t p) {
    int64_t r = -1;
    if (phip_binary_permanent_residues(A, n, &p, 1, &r) != 0) return -1;
    return r;
}

int phip_permanent_mod3(const uint64_t* mags, const uint64_t* sgns, int n) {
    DeviceContext* ctx = DeviceContext::current();
    if (!ctx) return -1;
    std::lock_guard<std::mutex> lock(ctx->mutex());
    if (phip::tuning().mod3_blocks && n >= phip::kBlockMinOrder && n <= phip::kBatchMaxOrder && mags && sgns) {
        // The modulo-3 block walk: the residues of the entries as -1, 0, 1
        // and the persistent engine's block walk with modulus 3 (row sums
        // that are multiples of 3 zero a term, so almost every state is
        // skipped 32 at a time).
        std::vector<int64_t> a((size_t)n * n, 0);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                if ((mags[j] >> i) & 1u) a[(size_t)i * n + j] = ((sgns[j] >> i) & 1u) ? -1 : 1;
        phip::PlanInput in{a.data(), n, 3u, 0};
        in.mod3 = true;
        int64_t r = -1;
        if (phip::run_plan_in_context(ctx, {in}, &r) != 0) return -1;
        return (int)r;
    }
    return phip::mod3_from_columns(ctx, mags, sgns, n);
}

int phip_permanent_mod3_matrix(const int* A, int n) {
    if (n < 0 || n >= 64) {
        phip::set_error("matrix size must be less than 64");
        return -1;
    }
    if (n == 0) return 1;
    if (!A) {
        phip::set_error("matrix pointer is null");
        return -1;
    }
    std::vector<uint64_t> mags(n, 0), sgns(n, 0);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            const int v = ((A[(size_t)i * n + j] % 3) + 3) % 3;
            if (v) {
                mags[j] |= 1ull << i;
                if (v == 2) sgns[j] |= 1ull << i;
            }
        }
    }
    return phip_permanent_mod3(mags.data(), sgns.data(), n);
}

}  // extern "C"
