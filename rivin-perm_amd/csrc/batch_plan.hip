// Host side of the persistent batch: plan construction (device-resident
// tables, cost-sorted work items), graph capture and replay, and the C ABI.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "batch.h"
#include "device_context.h"
#include "exact_limbs.h"
#include "launchers.h"
#include "permhip.h"
#include "phip_arith.h"
#include "phip_reduce.h"
#include "residues.h"
#include "tables.h"

namespace phip {

namespace {

__global__ void batch_reset_kernel(uint32_t* counters, int n) {
    const int i = threadIdx.x;
    if (i < n) counters[i] = 0;
}

// One workgroup per requested residue: fold the item sums of the slot's job
// for its prime, reduce, scale.
__global__ __launch_bounds__(256) void batch_finalize_kernel(const BatchJob* __restrict__ jobs,
                                                             const BatchSlot* __restrict__ slots,
                                                             const uint64_t* __restrict__ item_partials,
                                                             int64_t* __restrict__ residues) {
    const BatchSlot sl = slots[blockIdx.x];
    const BatchJob jb = jobs[sl.job];
    const uint64_t p = sl.p;
    if (is_exact_variant(jb.variant)) {
        // The job's items hold jb.set limb columns of exact signed sums,
        // |column| < 2^(37 + cbits) per item (cbits <= 24), so runs of
        // 2^(24 - cbits) items add exactly in int64 (< 2^61) before each
        // reduction.  Column c weighs 2^(31 c).
        const int ncols = jb.set;
        const uint32_t run_len = 1u << (jb.cbits < 24 ? 24 - jb.cbits : 0);
        int64_t run[kExactMaxColumns] = {};
        uint64_t col[kExactMaxColumns] = {};
        uint32_t in_run = 0;
        const uint64_t* parts = item_partials + (size_t)jb.first_item * kItemSlots;
        // Four items per pass with every load issued first: the item sums
        // come from L2 and one item per pass left each thread waiting on a
        // load per item.
        constexpr int kAhead = 4;
        for (uint32_t i0 = threadIdx.x; i0 < jb.n_items; i0 += 256 * kAhead) {
            int64_t v[kAhead][kExactMaxColumns];
#pragma unroll
            for (int u = 0; u < kAhead; ++u) {
                const uint32_t i = i0 + 256u * u;
#pragma unroll
                for (int c = 0; c < kExactMaxColumns; ++c)
                    v[u][c] = (i < jb.n_items && c < ncols) ? (int64_t)parts[(size_t)i * kItemSlots + c] : 0;
            }
#pragma unroll
            for (int u = 0; u < kAhead; ++u) {
                if (i0 + 256u * u >= jb.n_items) break;
#pragma unroll
                for (int c = 0; c < kExactMaxColumns; ++c) run[c] += v[u][c];
                if (++in_run == run_len) {
#pragma unroll
                    for (int c = 0; c < kExactMaxColumns; ++c) {
                        const int64_t r = run[c] % (int64_t)p;
                        col[c] += (uint64_t)(r < 0 ? r + (int64_t)p : r);
                        run[c] = 0;
                    }
                    in_run = 0;
                }
            }
        }
        uint64_t s = 0, w = 1 % p;
        const uint64_t step = (1ull << 31) % p;
#pragma unroll
        for (int c = 0; c < kExactMaxColumns; ++c) {
            const int64_t v = run[c] % (int64_t)p;
            const uint64_t r = (col[c] + (uint64_t)(v < 0 ? v + (int64_t)p : v)) % p;
            s += (r * w) % p;
            w = (w * step) % p;
        }
        s = block_reduce_sum_u64<256>(s % p);
        if (threadIdx.x == 0) residues[blockIdx.x] = (int64_t)(((s % p) * (uint64_t)sl.scale) % p);
        return;
    }
    // Item sums are signed (two's complement) sums of 64 lane residues,
    // |sum| < 2^37, so 2^16 of them add exactly in int64 (< 2^53); each such
    // run is reduced once instead of every item.
    uint64_t s = 0;
    int64_t run = 0;
    uint32_t in_run = 0;
    for (uint32_t i = threadIdx.x; i < jb.n_items; i += 256) {
        run += (int64_t)item_partials[(size_t)(jb.first_item + i) * kItemSlots + sl.k];
        if (++in_run == 65536u) {
            const int64_t v = run % (int64_t)p;
            s += (uint64_t)(v < 0 ? v + (int64_t)p : v);
            run = 0;
            in_run = 0;
        }
    }
    const int64_t v = run % (int64_t)p;
    s += (uint64_t)(v < 0 ? v + (int64_t)p : v);
    s = block_reduce_sum_u64<256>(s);
    if (threadIdx.x == 0) residues[blockIdx.x] = (int64_t)(((s % p) * (uint64_t)sl.scale) % p);
}

BatchLaunchFn launcher_for(int group, int set, int variant) {
#define PHIP_LAUNCHER(G, K) \
    if (group == G && set == K && variant == kVariantPlain) return launch_batch_g##G##_k##K;
#define PHIP_LAUNCHER_PAIRED(G, K) \
    if (group == G && set == K && variant == kVariantPaired) return launch_batch_g##G##_k##K##_paired;
#define PHIP_LAUNCHER_SKIP(G, K) \
    if (group == G && set == K && variant == kVariantSkipZeros) return launch_batch_g##G##_k##K##_skip;
    PHIP_LAUNCHER(0, 1)
    PHIP_LAUNCHER(2, 1) PHIP_LAUNCHER(2, 2) PHIP_LAUNCHER(2, 4) PHIP_LAUNCHER(2, 8)
    PHIP_LAUNCHER(3, 1) PHIP_LAUNCHER(3, 2) PHIP_LAUNCHER(3, 4) PHIP_LAUNCHER(3, 8)
    PHIP_LAUNCHER(4, 1) PHIP_LAUNCHER(4, 2) PHIP_LAUNCHER(4, 4) PHIP_LAUNCHER(4, 8)
    PHIP_LAUNCHER(6, 1) PHIP_LAUNCHER(6, 2) PHIP_LAUNCHER(6, 4) PHIP_LAUNCHER(6, 8)
    PHIP_LAUNCHER_PAIRED(2, 4) PHIP_LAUNCHER_PAIRED(2, 8) PHIP_LAUNCHER_PAIRED(3, 4) PHIP_LAUNCHER_PAIRED(3, 8)
    PHIP_LAUNCHER_PAIRED(4, 4) PHIP_LAUNCHER_PAIRED(4, 8) PHIP_LAUNCHER_PAIRED(6, 4) PHIP_LAUNCHER_PAIRED(6, 8)
    PHIP_LAUNCHER_SKIP(4, 2) PHIP_LAUNCHER_SKIP(4, 4) PHIP_LAUNCHER_SKIP(4, 8)
    PHIP_LAUNCHER_SKIP(6, 2) PHIP_LAUNCHER_SKIP(6, 4) PHIP_LAUNCHER_SKIP(6, 8)
    if (variant == kVariantExact) {
        if (group == 2 && set == 8) return launch_batch_g2_k8_exact;
        if (group == 3 && set == 8) return launch_batch_g3_k8_exact;
        if (group == 4 && set == 8) return launch_batch_g4_k8_exact;
        if (group == 6 && set == 8) return launch_batch_g6_k8_exact;
        if (group == 2 && set == 10) return launch_batch_g2_k10_exact;
        if (group == 2 && set == 12) return launch_batch_g2_k12_exact;
        if (group == 3 && set == 10) return launch_batch_g3_k10_exact;
        if (group == 3 && set == 12) return launch_batch_g3_k12_exact;
        if (group == 4 && set == 10) return launch_batch_g4_k10_exact;
        if (group == 4 && set == 12) return launch_batch_g4_k12_exact;
    }
    if (variant == kVariantExactCompact && group == 2 && set == 10) return launch_batch_g2_k10_compact;
    if (variant == kVariantSkipExact && group == 6 && set == 8) return launch_batch_g6_k8_skipexact;
    if (variant == kVariantExactSigned && group == 3 && set == 8) return launch_batch_g3_k8_signed;
    if (variant == kVariantExactSigned && group == 4 && set == 8) return launch_batch_g4_k8_signed;
#undef PHIP_LAUNCHER
#undef PHIP_LAUNCHER_PAIRED
#undef PHIP_LAUNCHER_SKIP
    return nullptr;
}

int occupancy_for(int group, int set, int variant, unsigned lds) {
#define PHIP_OCC(G, K) \
    if (group == G && set == K && variant == kVariantPlain) return occupancy_batch_g##G##_k##K(lds);
#define PHIP_OCC_PAIRED(G, K) \
    if (group == G && set == K && variant == kVariantPaired) return occupancy_batch_g##G##_k##K##_paired(lds);
#define PHIP_OCC_SKIP(G, K) \
    if (group == G && set == K && variant == kVariantSkipZeros) return occupancy_batch_g##G##_k##K##_skip(lds);
    PHIP_OCC(0, 1)
    PHIP_OCC(2, 1) PHIP_OCC(2, 2) PHIP_OCC(2, 4) PHIP_OCC(2, 8)
    PHIP_OCC(3, 1) PHIP_OCC(3, 2) PHIP_OCC(3, 4) PHIP_OCC(3, 8)
    PHIP_OCC(4, 1) PHIP_OCC(4, 2) PHIP_OCC(4, 4) PHIP_OCC(4, 8)
    PHIP_OCC(6, 1) PHIP_OCC(6, 2) PHIP_OCC(6, 4) PHIP_OCC(6, 8)
    PHIP_OCC_PAIRED(2, 4) PHIP_OCC_PAIRED(2, 8) PHIP_OCC_PAIRED(3, 4) PHIP_OCC_PAIRED(3, 8)
    PHIP_OCC_PAIRED(4, 4) PHIP_OCC_PAIRED(4, 8) PHIP_OCC_PAIRED(6, 4) PHIP_OCC_PAIRED(6, 8)
    PHIP_OCC_SKIP(4, 2) PHIP_OCC_SKIP(4, 4) PHIP_OCC_SKIP(4, 8)
    PHIP_OCC_SKIP(6, 2) PHIP_OCC_SKIP(6, 4) PHIP_OCC_SKIP(6, 8)
    if (variant == kVariantExact) {
        if (group == 2 && set == 8) return occupancy_batch_g2_k8_exact(lds);
        if (group == 3 && set == 8) return occupancy_batch_g3_k8_exact(lds);
        if (group == 4 && set == 8) return occupancy_batch_g4_k8_exact(lds);
        if (group == 6 && set == 8) return occupancy_batch_g6_k8_exact(lds);
        if (group == 2 && set == 10) return occupancy_batch_g2_k10_exact(lds);
        if (group == 2 && set == 12) return occupancy_batch_g2_k12_exact(lds);
        if (group == 3 && set == 10) return occupancy_batch_g3_k10_exact(lds);
        if (group == 3 && set == 12) return occupancy_batch_g3_k12_exact(lds);
        if (group == 4 && set == 10) return occupancy_batch_g4_k10_exact(lds);
        if (group == 4 && set == 12) return occupancy_batch_g4_k12_exact(lds);
    }
    if (variant == kVariantExactCompact && group == 2 && set == 10) return occupancy_batch_g2_k10_compact(lds);
    if (variant == kVariantSkipExact && group == 6 && set == 8) return occupancy_batch_g6_k8_skipexact(lds);
    if (variant == kVariantExactSigned && group == 3 && set == 8) return occupancy_batch_g3_k8_signed(lds);
    if (variant == kVariantExactSigned && group == 4 && set == 8) return occupancy_batch_g4_k8_signed(lds);
#undef PHIP_OCC
#undef PHIP_OCC_PAIRED
#undef PHIP_OCC_SKIP
    return -1;
}

// Sets of 4 or 8 primes all below kPairedPrimeLimit run on the paired-state
// kernels.
bool pairs_states(int set, const uint32_t* primes) {
    if (set != 4 && set != 8) return false;
    for (int k = 0; k < set; ++k)
        if (primes[k] >= kPairedPrimeLimit) return false;
    return true;
}

// Set sizes a table's r moduli are split into.  A set of K primes costs one
// shared walk (flips, exact group products) plus K per-prime Montgomery
// trees; the shared part is worth about two trees, so five primes run
// cheaper as 4 + 1 than as 8 with three padded slots, and six or seven as
// one padded set of 8.
std::vector<int> split_set(int r) {
    std::vector<int> out;
    while (r > 0) {
        int k;
        if (r >= 6) {
            k = 8;
        } else if (r == 5 || r == 4 || r == 3) {
            k = 4;
        } else {
            k = r;  // 2 or 1
        }
        out.push_back(k);
        r -= std::min(r, k);
    }
    return out;
}

// Sets of a 0/1 table on the zero-skipping kernels: its walk costs far more
// than the per-prime trees, which run on the few nonzero states only, so
// up to eight primes share one walk as one set of 2, 4 or 8 (padded).
std::vector<int> split_skip_set(int r) {
    std::vector<int> out;
    while (r > 0) {
        const int k = r <= 2 ? 2 : (r <= 4 ? 4 : 8);
        out.push_back(k);
        r -= std::min(r, k);
    }
    return out;
}

int groups_of(int n, int group) { return group == 0 ? n : (n + group - 1) / group; }

// Bit length of the product of the row bounds (0 when a bound is 0): every
// state's product of row sums, and so of group products, is at most that
// product in magnitude.
int bound_bits(const std::vector<uint64_t>& S) {
    std::vector<uint32_t> x(1, 1);  // little-endian 32-bit words
    for (uint64_t v : S) {
        if (v == 0) return 0;
        const uint64_t lo = v & 0xffffffffu, hi = v >> 32;
        std::vector<uint32_t> y(x.size() + 2, 0);
        for (int part = 0; part < 2; ++part) {
            const uint64_t f = part ? hi : lo;
            if (!f) continue;
            uint64_t carry = 0;
            size_t i = 0;
            for (; i < x.size(); ++i) {
                const unsigned __int128 t = (unsigned __int128)x[i] * f + y[i + part] + carry;
                y[i + part] = (uint32_t)t;
                carry = (uint64_t)(t >> 32);
            }
            for (size_t j = i + part; carry; ++j) {
                const uint64_t t = (uint64_t)y[j] + carry;
                y[j] = (uint32_t)t;
                carry = t >> 32;
            }
        }
        while (y.size() > 1 && y.back() == 0) y.pop_back();
        x.swap(y);
    }
    int bits = 32 * (int)(x.size() - 1);
    for (uint32_t top = x.back(); top; top >>= 1) ++bits;
    return bits;
}

// Limb columns of the exact variant for a packed table, or 0 when the table
// stays on the modular kernels: 8 for at most eight groups (full products),
// 10 or 12 for 9..16 groups whose row-bound product is below 2^310 or 2^372
// (products kept modulo 2^(31 L)).  `primes` is the number of moduli the
// walk would serve.
int exact_columns(int groups, int bits, size_t primes) {
    const Tuning& t = tuning();
    if (t.exact_min_primes <= 0) return 0;
    if (groups <= kExactMaxGroups) return primes >= (size_t)t.exact_min_primes ? kExactColumns : 0;
    if (groups > kExactWideMaxGroups || t.exact_wide_min_primes <= 0 || primes < (size_t)t.exact_wide_min_primes)
        return 0;
    if (bits <= 31 * 10) return 10;
    if (bits <= 31 * kExactMaxColumns) return kExactMaxColumns;
    return 0;
}

// Whether a table of groups of G rows (row bounds `S` in kernel row order)
// fits the compact exact classes: every node [lo, hi) of the product tree
// (LimbProduct<0, groups, L, kCompactValueBits>: split at limb_split) whose
// limb count is below L holds its exact product, i.e. the product of its
// rows' bounds has at most 31 limbs_for(hi - lo, V) bits.  Nodes of L limbs
// are products modulo 2^(31 L), exact at the root by the table's bound.
bool compact_nodes_fit(const std::vector<uint64_t>& S, int G, int lo, int hi, int L) {
    const int count = hi - lo;
    if (count <= 1) return true;
    const int limbs = limbs_for(count, kCompactValueBits);
    if (limbs < L) {
        const int n = (int)S.size();
        const std::vector<uint64_t> rows(S.begin() + lo * G, S.begin() + std::min(n, hi * G));
        if (bound_bits(rows) > 31 * limbs) return false;
    }
    const int mid = lo + limb_split(count);
    return compact_nodes_fit(S, G, lo, mid, L) && compact_nodes_fit(S, G, mid, hi, L);
}

// Whether a table of at most eight groups of G rows (row bounds S in kernel
// order) fits the signed-limb exact classes: the product of the row bounds
// of groups [0, 4) and of groups [4, 8) each below 2^kSignedQuadBits.
bool signed_quads_fit(const std::vector<uint64_t>& S, int G) {
    const int n = (int)S.size();
    for (int lo = 0; lo < n; lo += 4 * G) {
        const std::vector<uint64_t> rows(S.begin() + lo, S.begin() + std::min(n, lo + 4 * G));
        if (bound_bits(rows) > kSignedQuadBits) return false;
    }
    return true;
}

// Estimated share of the Gray states whose term is nonzero, for a 0/1
// matrix: the product over rows of 1 - C(d, d/2) / 2^d, the share of sign
// vectors that do not zero a row of even degree d (with or without the
// fixed column in the row; a row of odd degree never sums to zero).  Rows
// are treated as independent.  1 for any other matrix.
double nonzero_share(const int64_t* a, int n) {
    double share = 1.0;
    for (int i = 0; i < n; ++i) {
        int d = 0;
        for (int j = 0; j < n; ++j) {
            const int64_t v = a[(size_t)i * n + j];
            if (v != 0 && v != 1) return 1.0;
            d += (int)v;
        }
        if (d % 2 == 0) {
            double z = 1.0;
            for (int t = 1; t <= d / 2; ++t) z *= (2.0 * t - 1.0) / (2.0 * t);
            share *= 1.0 - z;
        }
    }
    return share;
}

// Row order and zero-test width of a 0/1 table on the zero-skipping walk.
// A row of odd degree sums to an odd number in every state, so only rows of
// even degree can zero a term.  Rows of even degree go first, smallest
// degree first (the likeliest to sum to zero), then the rows of odd degree;
// when every even row lies in the leading skip_half_words(n) words and each
// group of `group` consecutive rows of that order can be arranged to meet
// the kernel's operand limits (group_fits on the degrees, which are the
// row bounds of a 0/1 matrix), `order` becomes that order and the test
// reads only those words.  Otherwise `order` is left as it is (plan_rows'
// order) and the test reads every word.  Returns the words tested.
int skip_plan(const int64_t* a, int n, int group, std::vector<int>& order) {
    std::vector<uint64_t> degree(n, 0);
    std::vector<int> even, odd;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) degree[i] += a[(size_t)i * n + j] != 0 ? 1u : 0u;
        (degree[i] % 2 == 0 ? even : odd).push_back(i);
    }
    const int half_rows = 4 * skip_half_words(n);
    if ((int)even.size() > half_rows || skip_half_words(n) == skip_full_words(n)) return skip_full_words(n);
    std::stable_sort(even.begin(), even.end(), [&](int x, int y) { return degree[x] < degree[y]; });
    std::vector<int> cand(even);
    cand.insert(cand.end(), odd.begin(), odd.end());
    // Arrangements within a group are tried from the even-first one on;
    // rows of equal degree are interchangeable for group_fits.
    auto key = [&](int x, int y) {
        return std::make_pair(degree[x] % 2, degree[x]) < std::make_pair(degree[y] % 2, degree[y]);
    };
    for (int first = 0; first < n; first += group) {
        const auto lo = cand.begin() + first;
        const auto hi = cand.begin() + std::min(n, first + group);
        std::stable_sort(lo, hi, key);
        bool fits = false;
        do {
            uint64_t b[6];
            for (auto it = lo; it != hi; ++it) b[it - lo] = degree[*it];
            if (group_fits(b, (int)(hi - lo))) {
                fits = true;
                break;
            }
        } while (std::next_permutation(lo, hi, key));
        if (!fits) return skip_full_words(n);
    }
    for (int pos = half_rows; pos < n; ++pos)
        if (degree[cand[pos]] % 2 == 0) return skip_full_words(n);
    order = cand;
    return skip_half_words(n);
}

// Row order and row-sum group of a 0/1 table on the block zero-skipping
// walk (build_block_tables): the rows of even degree first, since the
// block test reads the leading rows, then the rows of odd degree, each by
// ascending degree.  The row bounds are the degrees d: `group` becomes 6
// when every group of six consecutive rows in this order meets the
// operand limits, else 4, which every order meets for degrees up to 46
// (46^4 < 2^31, 46^2 < 2^23).  Returns the number of rows of even degree,
// or -1 when neither fits.
int block_plan(const int64_t* a, int n, std::vector<int>& order, int& group) {
    std::vector<uint64_t> degree(n, 0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) degree[i] += a[(size_t)i * n + j] != 0 ? 1u : 0u;
    order.resize(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return std::make_pair(degree[x] % 2, degree[x]) < std::make_pair(degree[y] % 2, degree[y]);
    });
    int even = 0;
    while (even < n && degree[order[even]] % 2 == 0) ++even;
    auto fits = [&](int g) {
        for (int first = 0; first < n; first += g) {
            uint64_t b[6];
            const int c = std::min(g, n - first);
            for (int k = 0; k < c; ++k) b[k] = degree[order[first + k]];
            if (!group_fits(b, c)) return false;
        }
        return true;
    };
    if (fits(6)) {
        group = 6;
    } else if (fits(4)) {
        group = 4;
    } else {
        return -1;
    }
    return even;
}

template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    ~DevBuf() { release(); }
    void release() {
        if (ptr) hipFree(ptr);
        ptr = nullptr;
    }
    int alloc(size_t count) {
        if (ptr) {
            hipFree(ptr);
            ptr = nullptr;
        }
        hipError_t e = hipMalloc(&ptr, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) {
            set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
            ptr = nullptr;
            return -1;
        }
        return 0;
    }
};

}  // namespace

int build_plan_host(const std::vector<PlanInput>& in, int items_log2, int min_cbits, PlanHost& out) {
    out = PlanHost();
    // Pass 1: validate; one table per distinct (key, order, representation,
    // table contents); the inputs that share it form a set.
    struct Set {
        int key;
        int n;
        int group;
        uint64_t data_off;
        bool skip_zeros;
        bool compact;    // groups of 2 rows whose partial bound products fit the compact exact classes
        bool signed_ok;  // groups of 3 or 4 rows whose four-group bound products fit the signed-limb classes
        int test_words;  // skip_zeros: the walk's zero-test width (skip_plan)
        int block_rows;  // skip_zeros: rows the block walk's test reads, or -1 (per-state test)
        bool mod3;       // the modulo-3 block walk
        int bits;  // bound_bits of the table's row bounds
        std::vector<size_t> inputs;
    };
    std::vector<Set> sets;
    std::map<int, std::vector<size_t>> sets_of_key;
    std::vector<int32_t> table;
    // Every modulus above twice a matrix's largest |entry| sees the matrix
    // itself as its centred table.  The first such input of a matrix plans
    // the table; later ones join its set without planning it again.
    struct Plain {
        uint64_t maxabs;
        size_t set;
    };
    std::map<std::tuple<int, const int64_t*, int>, Plain> plain_of;
    for (size_t i = 0; i < in.size(); ++i) {
        const PlanInput& x = in[i];
        if (x.n < kBatchMinOrder || x.n > kBatchMaxOrder) {
            set_error("batch orders must lie in [5, 46]; use the single-matrix entry point otherwise");
            return -1;
        }
        if (x.q < 3 || x.q >= (1u << 31) || !(x.q & 1)) {
            set_error("batch moduli must be odd and lie in [3, 2^31)");
            return -1;
        }
        const auto pkey = std::make_tuple(x.key, x.a, x.n);
        auto pit = plain_of.find(pkey);
        if (pit == plain_of.end()) {
            uint64_t m = 0;
            for (size_t e = 0; e < (size_t)x.n * x.n; ++e) {
                const int64_t v = x.a[e];
                m = std::max(m, v < 0 ? 0 - (uint64_t)v : (uint64_t)v);
            }
            pit = plain_of.emplace(pkey, Plain{m, SIZE_MAX}).first;
        }
        if (x.mod3 && (x.q != 3 || pit->second.maxabs > 1 || x.n < kBlockMinOrder)) {
            set_error("internal: a modulo-3 block walk needs entries in {-1, 0, 1}, modulus 3 and order >= 8");
            return -1;
        }
        const bool plain = pit->second.maxabs <= (uint64_t)(x.q - 1) / 2;
        if (plain && pit->second.set != SIZE_MAX && sets[pit->second.set].mod3 == x.mod3) {
            sets[pit->second.set].inputs.push_back(i);
            continue;
        }
        // A 0/1 matrix walks in the orientation (it or its transpose, the
        // same permanent) with the smaller estimated share of nonzero terms:
        // only lines of even degree can sum to zero, and the counts of even
        // rows and of even columns differ by chance.  The taken states are
        // most of a 0/1 walk's work (about 12% fewer at the served orders
        // 35-38).
        const int64_t* a = x.a;
        std::vector<int64_t> transposed;
        double share = nonzero_share(x.a, x.n);
        bool binary = !x.mod3;
        for (size_t e = 0; binary && e < (size_t)x.n * x.n; ++e) binary = x.a[e] == 0 || x.a[e] == 1;
        if (binary) {
            transposed.resize((size_t)x.n * x.n);
            for (int r = 0; r < x.n; ++r)
                for (int c = 0; c < x.n; ++c) transposed[(size_t)c * x.n + r] = x.a[(size_t)r * x.n + c];
            const double share_t = nonzero_share(transposed.data(), x.n);
            if (share_t < share) {
                a = transposed.data();
                share = share_t;
            }
        }
        std::vector<int> order;
        const std::vector<uint64_t> bounds = centred_row_bounds(a, x.n, (int64_t)x.q);
        int group = plan_rows(bounds, order);
        // Tables of 0/1 matrices (nonzero_share is 1 for any other entry) in
        // groups of 4 or 6 with a small share of nonzero terms run on the
        // zero-skipping kernels, whose walk reads byte tables in the same
        // row order.
        const int permille = tuning().skip_zeros_permille;
        const bool skip = x.mod3 || (permille > 0 && (group == 4 || group == 6) &&
                                     share * 1000.0 <= (double)permille);
        // Those of order kBlockMinOrder and above walk in blocks of 32 states
        // (build_block_tables, block_plan's row order and group).
        int test_words = 0, block_rows = -1, even = -1;
        if (x.mod3) {
            // Every row may sum to a multiple of 3: all of them are tested.
            even = block_plan(a, x.n, order, group);
            if (even < 0) {
                set_error("internal: no row-sum group fits a modulo-3 table");
                return -1;
            }
            block_rows = block_test_rows(x.n, x.n);
        } else if (skip && tuning().skip_blocks && x.n >= kBlockMinOrder) {
            std::vector<int> block_order;
            int block_group = group;
            even = block_plan(a, x.n, block_order, block_group);
            if (even >= 0) {
                order = block_order;
                group = block_group;
                block_rows = block_test_rows(x.n, even);
            }
        }
        size_t words;
        if (x.mod3) {
            words = block_table_words_mod3(x.n, block_rows);
        } else if (block_rows >= 0) {
            words = block_table_words(x.n, block_rows);
        } else {
            words = (size_t)x.n * (skip ? byte_stride(x.n) : glynn_stride(x.n, group));
            test_words = skip ? skip_plan(a, x.n, group, order) : 0;
        }
        table.assign(words, 0);
        if (x.mod3) {
            build_block_tables_mod3(a, x.n, order, block_rows, table.data());
        } else if (block_rows >= 0) {
            build_block_tables(a, x.n, order, even, block_rows, table.data());
        } else if (skip) {
            build_byte_tables(a, x.n, order, table.data());
        } else {
            build_glynn_tables(a, x.n, x.q, group, order, table.data());
        }
        size_t found = sets.size();
        if (group > 0) {
            for (size_t si : sets_of_key[x.key]) {
                const Set& st = sets[si];
                if (st.n == x.n && st.group == group && st.skip_zeros == skip && st.test_words == test_words &&
                    st.block_rows == block_rows && st.mod3 == x.mod3 &&
                    memcmp(out.blob.data() + st.data_off, table.data(), words * sizeof(int32_t)) == 0) {
                    found = si;
                    break;
                }
            }
        }
        if (found == sets.size()) {
            bool compact = false;
            if (group == 2 && !skip && tuning().exact_compact) {
                const int groups = groups_of(x.n, group);
                if (groups > kExactMaxGroups && groups <= kExactWideMaxGroups) {
                    std::vector<uint64_t> kernel_bounds(x.n);
                    for (int r = 0; r < x.n; ++r) kernel_bounds[r] = bounds[order.empty() ? r : order[r]];
                    compact = compact_nodes_fit(kernel_bounds, group, 0, groups, 10);
                }
            }
            bool signed_ok = false;
            if ((group == 3 || group == 4) && !skip && tuning().exact_signed &&
                groups_of(x.n, group) <= kExactMaxGroups) {
                std::vector<uint64_t> kernel_bounds(x.n);
                for (int r = 0; r < x.n; ++r) kernel_bounds[r] = bounds[order.empty() ? r : order[r]];
                signed_ok = signed_quads_fit(kernel_bounds, group);
            }
            sets.push_back({x.key, x.n, group, (uint64_t)out.blob.size(), skip, compact, signed_ok, test_words, block_rows,
                            x.mod3, group > 0 ? bound_bits(bounds) : 0, {}});
            sets_of_key[x.key].push_back(found);
            out.blob.insert(out.blob.end(), table.begin(), table.end());
        }
        sets[found].inputs.push_back(i);
        if (plain && group > 0) pit->second.set = found;
    }

    // Pass 2: jobs (one per set chunk) and slots.
    out.slots.resize(in.size());
    for (const Set& st : sets) {
        size_t next = 0;
        // A packed table with enough primes runs one exact-accumulation
        // walk for all of them (exact_columns): with eight groups the walk
        // costs about as much as the modular walk of a set of six primes,
        // and the residues come from its exact sums.
        // A 0/1 block table of at most eight groups of 6 rows with enough
        // primes runs the block walk once with exact accumulation (a taken
        // state costs its limb product instead of a Montgomery tree per
        // prime).  Block tables fall back to groups of 4 only above order
        // 35 (group_fits on degrees up to 35 holds for 6), i.e. with more
        // than eight groups.
        const bool skip_exact = st.skip_zeros && st.block_rows >= 0 && !st.mod3 && st.group == 6 &&
                                groups_of(st.n, st.group) <= kExactMaxGroups && tuning().skip_exact_min_primes > 0 &&
                                st.inputs.size() >= (size_t)tuning().skip_exact_min_primes;
        const int columns = skip_exact ? kExactColumns
                            : (st.group > 0 && !st.skip_zeros)
                                ? exact_columns(groups_of(st.n, st.group), st.bits, st.inputs.size())
                                : 0;
        if (columns > 0) {
            BatchJob jb{};
            jb.data_off = st.data_off;
            jb.n = st.n;
            jb.group = st.group;
            jb.set = columns;
            jb.nprimes = (int32_t)st.inputs.size();
            jb.prime_off = (uint32_t)out.set_p.size();
            jb.variant = skip_exact                     ? kVariantSkipExact
                         : (columns == 10 && st.compact) ? kVariantExactCompact
                         : (columns == 8 && st.signed_ok) ? kVariantExactSigned
                                                          : kVariantExact;
            jb.block_rows = skip_exact ? st.block_rows : -1;
            jb.test_words = skip_exact ? st.test_words : 0;
            jb.block_mod3 = 0;
            const uint32_t job = (uint32_t)out.jobs.size();
            const int L = st.n - 1;
            for (size_t k = 0; k < st.inputs.size(); ++k) {
                const size_t idx = st.inputs[k];
                const uint32_t q = in[idx].q;
                BatchSlot sl;
                sl.job = job;
                sl.k = (uint32_t)k;
                sl.p = q;
                // The exact sums carry only Glynn's 2^(n-1).
                sl.scale = (uint32_t)host_powmod((q + 1) / 2, (uint64_t)L, q);
                out.slots[idx] = sl;
                out.total_states += std::ldexp(1.0, L);
                out.total_mulmods += (double)st.n * std::ldexp(1.0, L);
            }
            out.jobs.push_back(jb);
            continue;
        }
        std::vector<int> set_sizes =
            st.group == 0 ? std::vector<int>(st.inputs.size(), 1)
                          : (st.skip_zeros ? split_skip_set((int)st.inputs.size()) : split_set((int)st.inputs.size()));
        for (int K : set_sizes) {
            const int take = (int)std::min<size_t>((size_t)K, st.inputs.size() - next);
            BatchJob jb{};
            jb.data_off = st.data_off;
            jb.n = st.n;
            jb.group = st.group;
            jb.set = K;
            jb.nprimes = take;
            jb.prime_off = (uint32_t)out.set_p.size();
            const uint32_t job = (uint32_t)out.jobs.size();
            const int L = st.n - 1;
            const int groups = groups_of(st.n, st.group);
            for (int k = 0; k < K; ++k) {
                const size_t idx = st.inputs[next + std::min(k, take - 1)];
                const uint32_t q = in[idx].q;
                out.set_p.push_back((int32_t)q);
                out.set_pinv.push_back(inverse_mod_2_32(q));
                if (k < take) {
                    BatchSlot sl;
                    sl.job = job;
                    sl.k = (uint32_t)k;
                    sl.p = q;
                    // The tree leaves R^-(groups-1) and the lane reduction one
                    // more R^-1.
                    sl.scale = (uint32_t)host_mulmod(host_powmod((1ull << 32) % q, (uint64_t)groups, q),
                                                     host_powmod((q + 1) / 2, (uint64_t)L, q), q);
                    out.slots[idx] = sl;
                    out.total_states += std::ldexp(1.0, L);
                    out.total_mulmods += (double)st.n * std::ldexp(1.0, L);
                }
            }
            next += take;
            jb.block_rows = -1;
            jb.block_mod3 = 0;
            if (st.skip_zeros) {
                jb.variant = kVariantSkipZeros;
                jb.test_words = st.test_words;
                jb.block_rows = st.block_rows;
                jb.block_mod3 = st.mod3 ? 1 : 0;
            } else if (st.group > 0 && pairs_states(K, (const uint32_t*)out.set_p.data() + jb.prime_off)) {
                jb.variant = kVariantPaired;
            } else {
                jb.variant = kVariantPlain;
            }
            out.jobs.push_back(jb);
        }
    }

    return assign_items(out, items_log2, min_cbits);
}

int assign_items(PlanHost& out, int items_log2, int min_cbits) {
    out.items.clear();
    out.classes.clear();
    out.items_log2 = items_log2;
    out.min_cbits = min_cbits;
    // Chunk bits: spread all job states over about 2^items_log2 items, and
    // each class's states over at least about 2^(items_log2 - 2).  Classes
    // run one after another, so a class with few states would otherwise be
    // a handful of long items, each wave alone on its SIMD while the rest
    // of the device idles.  The second rule never lengthens a chunk.
    auto spread = [](double states, double items) {
        const int s = (int)std::lround(std::log2(std::max(1.0, states / items)));
        return std::min(std::max(s, 10), 30);
    };
    auto class_key = [](const BatchJob& jb) { return (jb.group * 32 + jb.set) * 8 + jb.variant; };
    double job_states = 0;
    std::map<int, double> class_states;
    for (const BatchJob& jb : out.jobs) {
        job_states += std::ldexp(1.0, jb.n - 1);
        class_states[class_key(jb)] += std::ldexp(1.0, jb.n - 1);
    }
    const int s_all = spread(job_states, std::ldexp(1.0, items_log2));
    for (BatchJob& jb : out.jobs) {
        const int L = jb.n - 1;
        const int s = std::min(s_all, spread(class_states[class_key(jb)], std::ldexp(1.0, items_log2 - 2)));
        int c = (L >= 10) ? std::max(4, std::min(s - 6, L - 6)) : 4;
        c = std::min(std::max(c, min_cbits), L);
        // Exact jobs keep a lane's columns below 2^55 and an item's below
        // 2^61 (the finalisation's run length relies on it).
        if (is_exact_variant(jb.variant)) c = std::min(c, 24);
        // The block walk enumerates the first kBlockInnerBits Gray columns
        // inside a chunk (n >= kBlockMinOrder leaves L >= 7).
        if (is_skip_variant(jb.variant) && jb.block_rows >= 0) c = std::max(c, kBlockMinChunkBits);
        jb.cbits = c;
        jb.lanes = 1ull << (L - c);
    }

    // Items: per (group, set, variant) class, jobs by descending cost, each
    // job's items contiguous.
    std::vector<bool> placed(out.jobs.size(), false);
    static const int kClasses[][3] = {{6, 8, 5}, {3, 8, 6}, {4, 8, 6},
                                      {2, 8, 3}, {3, 8, 3}, {4, 8, 3}, {6, 8, 3}, {2, 10, 4},
                                      {2, 10, 3}, {2, 12, 3}, {3, 10, 3}, {3, 12, 3}, {4, 10, 3}, {4, 12, 3},
                                      {0, 1, 0},
                                      {2, 8, 1}, {2, 8, 0}, {2, 4, 1}, {2, 4, 0}, {2, 2, 0}, {2, 1, 0},
                                      {3, 8, 1}, {3, 8, 0}, {3, 4, 1}, {3, 4, 0}, {3, 2, 0}, {3, 1, 0},
                                      {4, 8, 1}, {4, 8, 2}, {4, 8, 0}, {4, 4, 1}, {4, 4, 2}, {4, 4, 0},
                                      {4, 2, 2}, {4, 2, 0}, {4, 1, 0},
                                      {6, 8, 1}, {6, 8, 2}, {6, 8, 0}, {6, 4, 1}, {6, 4, 2}, {6, 4, 0},
                                      {6, 2, 2}, {6, 2, 0}, {6, 1, 0}};
    for (const auto& gk : kClasses) {
        std::vector<uint32_t> order;
        for (uint32_t j = 0; j < out.jobs.size(); ++j)
            if (out.jobs[j].group == gk[0] && out.jobs[j].set == gk[1] && out.jobs[j].variant == gk[2])
                order.push_back(j);
        if (order.empty()) continue;
        auto cost = [&](uint32_t j) {
            const BatchJob& b = out.jobs[j];
            return std::ldexp(1.0, b.cbits) * (double)std::min<uint64_t>(b.lanes, 64) *
                   (b.n + 3.0 * groups_of(b.n, b.group));
        };
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost(a) > cost(b); });
        BatchClass cl{gk[0], gk[1], gk[2], (uint32_t)out.items.size(), 0, 0, 0};
        for (uint32_t j : order) {
            BatchJob& b = out.jobs[j];
            cl.max_n = std::max(cl.max_n, b.n);
            if (is_skip_variant(b.variant) && b.block_rows >= 0)
                cl.max_block_words =
                    std::max(cl.max_block_words, b.block_mod3 ? block_region_words_mod3(b.n, b.block_rows)
                                                              : block_region_words(b.n, b.block_rows));
            const uint64_t waves = (b.lanes + 63) / 64;
            if (out.items.size() + waves >= 0xffffffffull) {
                set_error("batch has too many work items");
                return -1;
            }
            b.first_item = (uint32_t)out.items.size();
            b.n_items = (uint32_t)waves;
            placed[j] = true;
            for (uint64_t w = 0; w < waves; ++w) out.items.push_back(make_uint2(j, (uint32_t)w));
        }
        cl.n_items = (uint32_t)out.items.size() - cl.first_item;
        out.classes.push_back(cl);
    }
    // A job of a class missing from the list above would get no items and
    // its residues would read as zero.
    for (bool p : placed)
        if (!p) {
            set_error("internal: a batch job has no kernel class");
            return -1;
        }
    return 0;
}

// Workgroups per CU that a class's kernel keeps resident with its dynamic
// LDS, or -1 when the runtime cannot say.
int class_resident(const BatchClass& cl) {
    return occupancy_for(cl.group, cl.set, cl.variant, class_lds_bytes(cl));
}

int enqueue_plan(const PlanHost& h, const PlanDevice& d, hipStream_t stream, int cus, int blocks_per_cu,
                 hipEvent_t ev0, hipEvent_t ev1) {
    if (h.slots.empty()) return 0;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, stream, d.counters, (int)h.classes.size());
    PHIP_HIP_TRY(hipGetLastError());
    if (ev0) PHIP_HIP_TRY(hipEventRecord(ev0, stream));
    for (size_t c = 0; c < h.classes.size(); ++c) {
        const BatchClass& cl = h.classes[c];
        const BatchLaunchFn fn = launcher_for(cl.group, cl.set, cl.variant);
        if (!fn) {
            set_error("internal: no kernel for a batch class");
            return -1;
        }
        // The persistent grid holds at most as many workgroups per CU as
        // the class's kernel keeps resident: more would only queue behind
        // the resident ones and start once the work is gone.  The
        // zero-skipping kernels always run every resident workgroup: their
        // walk is short dependent chains of full-rate adds and logic, and
        // more waves per SIMD hide them (binary36 40 -> 36 ms from 4 to 7
        // workgroups per CU).
        const int resident = class_resident(cl);
        const int want = (is_skip_variant(cl.variant) && resident > 0) ? resident : blocks_per_cu;
        const int per_cu = resident > 0 ? std::min(want, resident) : want;
        PHIP_HIP_TRY(fn(dim3((unsigned)(cus * std::max(per_cu, 1))), class_lds_bytes(cl), stream, d.jobs,
                        d.items + cl.first_item, cl.n_items, d.counters + c, d.blob, d.set_p, d.set_pinv,
                        d.partials + (size_t)cl.first_item * kItemSlots));
    }
    if (ev1) PHIP_HIP_TRY(hipEventRecord(ev1, stream));
    hipLaunchKernelGGL(batch_finalize_kernel, dim3((unsigned)h.slots.size()), dim3(256), 0, stream, d.jobs,
                       d.slots, d.partials, d.residues);
    PHIP_HIP_TRY(hipGetLastError());
    return 0;
}

struct BatchPlan {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    PlanHost host;
    DevBuf<BatchJob> d_jobs;
    DevBuf<uint2> d_items;
    DevBuf<int32_t> d_blob;
    DevBuf<int32_t> d_set_p;
    DevBuf<uint32_t> d_set_pinv;
    DevBuf<BatchSlot> d_slots;
    DevBuf<uint64_t> d_partials;
    DevBuf<uint32_t> d_counters;
    DevBuf<int64_t> d_res;
    int64_t* h_res = nullptr;
    int blocks_per_cu = 2;
    int cus = 256;
    bool use_graph = true;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0;
    bool launched = false;

    ~BatchPlan() {
        if (stream) hipStreamSynchronize(stream);
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        if (ev0) hipEventDestroy(ev0);
        if (ev1) hipEventDestroy(ev1);
        if (h_res) hipHostFree(h_res);
        if (stream) hipStreamDestroy(stream);
    }

    PlanDevice device_image() const {
        return PlanDevice{d_jobs.ptr,  d_items.ptr,    d_blob.ptr,     d_set_p.ptr, d_set_pinv.ptr,
                          d_slots.ptr, d_partials.ptr, d_counters.ptr, d_res.ptr};
    }

    int upload() {
        const PlanHost& h = host;
        if (d_blob.alloc(h.blob.size()) || d_set_p.alloc(h.set_p.size()) || d_set_pinv.alloc(h.set_pinv.size()) ||
            d_slots.alloc(h.slots.size()) || d_counters.alloc(32) || d_res.alloc(h.slots.size()))
            return -1;
        PHIP_HIP_TRY(hipHostMalloc((void**)&h_res, std::max<size_t>(h.slots.size(), 1) * sizeof(int64_t),
                                   hipHostMallocDefault));
        PHIP_HIP_TRY(hipMemcpy(d_blob.ptr, h.blob.data(), h.blob.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        PHIP_HIP_TRY(hipMemcpy(d_set_p.ptr, h.set_p.data(), h.set_p.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        PHIP_HIP_TRY(hipMemcpy(d_set_pinv.ptr, h.set_pinv.data(), h.set_pinv.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice));
        PHIP_HIP_TRY(
            hipMemcpy(d_slots.ptr, h.slots.data(), h.slots.size() * sizeof(BatchSlot), hipMemcpyHostToDevice));
        return upload_items();
    }

    // Jobs, items and partials: the parts that depend on the chunk sizes.
    int upload_items() {
        const PlanHost& h = host;
        if (d_jobs.alloc(h.jobs.size()) || d_items.alloc(h.items.size()) || d_partials.alloc(h.items.size() * kItemSlots))
            return -1;
        PHIP_HIP_TRY(hipMemcpy(d_jobs.ptr, h.jobs.data(), h.jobs.size() * sizeof(BatchJob), hipMemcpyHostToDevice));
        PHIP_HIP_TRY(hipMemcpy(d_items.ptr, h.items.data(), h.items.size() * sizeof(uint2), hipMemcpyHostToDevice));
        return 0;
    }

    void drop_graph() {
        if (exec) {
            hipGraphExecDestroy(exec);
            exec = nullptr;
        }
        if (graph) {
            hipGraphDestroy(graph);
            graph = nullptr;
        }
    }

    // New chunk sizes for about 2^items_log2 work items.
    int resize_items(int items_log2) {
        PHIP_HIP_TRY(hipSetDevice(device));
        PHIP_HIP_TRY(hipStreamSynchronize(stream));
        const int previous = host.items_log2;
        if (assign_items(host, items_log2, host.min_cbits) != 0) {
            // The previous sizing succeeded before; restore it so the plan
            // stays runnable.
            const std::string why = last_error();
            assign_items(host, previous, host.min_cbits);
            set_error(why);
            return -1;
        }
        if (upload_items() != 0) {
            // A half-written work list must never run: without it every
            // later launch reports the failed re-sizing.
            const std::string why = last_error();
            d_jobs.release();
            d_items.release();
            d_partials.release();
            set_error(why);
            return -1;
        }
        return 0;
    }

    int enqueue(bool timed) {
        if (enqueue_plan(host, device_image(), stream, cus, blocks_per_cu, timed ? ev0 : nullptr,
                         timed ? ev1 : nullptr) != 0)
            return -1;
        PHIP_HIP_TRY(hipMemcpyAsync(h_res, d_res.ptr, host.slots.size() * sizeof(int64_t), hipMemcpyDeviceToHost,
                                    stream));
        return 0;
    }

    // Enqueue one run on the plan's stream without waiting.
    int launch_locked() {
        if (!d_jobs.ptr || !d_items.ptr || !d_partials.ptr) {
            set_error("batch plan has no device work list (a failed re-sizing); create it again");
            return -1;
        }
        PHIP_HIP_TRY(hipSetDevice(device));
        if (use_graph) {
            if (!exec) {
                PHIP_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                if (enqueue(false) != 0) {
                    hipGraph_t g = nullptr;
                    hipStreamEndCapture(stream, &g);
                    if (g) hipGraphDestroy(g);
                    return -1;
                }
                PHIP_HIP_TRY(hipStreamEndCapture(stream, &graph));
                PHIP_HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            }
            PHIP_HIP_TRY(hipEventRecord(ev0, stream));
            PHIP_HIP_TRY(hipGraphLaunch(exec, stream));
            PHIP_HIP_TRY(hipEventRecord(ev1, stream));
        } else {
            if (enqueue(true) != 0) return -1;
        }
        launched = true;
        return 0;
    }

    // 1: finished (residues copied to out when not null), 0: still running.
    int finish_locked(int64_t* out, bool block) {
        if (!launched) {
            set_error("batch plan has not been launched");
            return -1;
        }
        if (block) {
            PHIP_HIP_TRY(hipStreamSynchronize(stream));
        } else {
            const hipError_t e = hipStreamQuery(stream);
            if (e == hipErrorNotReady) return 0;
            PHIP_HIP_TRY(e);
        }
        launched = false;
        PHIP_HIP_TRY(hipEventElapsedTime(&last_ms, ev0, ev1));
        if (out) memcpy(out, h_res, host.slots.size() * sizeof(int64_t));
        return 1;
    }

    int run(int64_t* out) {
        std::lock_guard<std::mutex> lock(mu);
        if (launched && finish_locked(nullptr, true) < 0) return -1;
        if (launch_locked() != 0) return -1;
        return finish_locked(out, true) < 0 ? -1 : 0;
    }
};

int run_plan_in_context(DeviceContext* ctx, const std::vector<PlanInput>& in, int64_t* out) {
    PlanHost h;
    const Tuning& t = tuning();
    if (build_plan_host(in, std::max(0, t.lanes_log2 - 6), t.min_cbits, h) != 0) return -1;
    if (h.slots.empty()) return 0;
    // Device buffers of the context, slots 16..24.
    PlanDevice d;
    d.jobs = (BatchJob*)ctx->buffer(16, h.jobs.size() * sizeof(BatchJob));
    d.items = (uint2*)ctx->buffer(17, h.items.size() * sizeof(uint2));
    d.blob = (int32_t*)ctx->buffer(18, h.blob.size() * sizeof(int32_t));
    d.set_p = (int32_t*)ctx->buffer(19, h.set_p.size() * sizeof(int32_t));
    d.set_pinv = (uint32_t*)ctx->buffer(20, h.set_pinv.size() * sizeof(uint32_t));
    d.slots = (BatchSlot*)ctx->buffer(21, h.slots.size() * sizeof(BatchSlot));
    d.partials = (uint64_t*)ctx->buffer(22, h.items.size() * kItemSlots * sizeof(uint64_t));
    d.counters = (uint32_t*)ctx->buffer(23, 32 * sizeof(uint32_t));
    d.residues = (int64_t*)ctx->buffer(24, h.slots.size() * sizeof(int64_t));
    auto* h_res = (int64_t*)ctx->host_buffer(4, h.slots.size() * sizeof(int64_t));
    if (!d.jobs || !d.items || !d.blob || !d.set_p || !d.set_pinv || !d.slots || !d.partials || !d.counters ||
        !d.residues || !h_res)
        return -1;
    hipStream_t st = ctx->stream();
    PHIP_HIP_TRY(hipMemcpyAsync(d.jobs, h.jobs.data(), h.jobs.size() * sizeof(BatchJob), hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d.items, h.items.data(), h.items.size() * sizeof(uint2), hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d.blob, h.blob.data(), h.blob.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(
        hipMemcpyAsync(d.set_p, h.set_p.data(), h.set_p.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d.set_pinv, h.set_pinv.data(), h.set_pinv.size() * sizeof(uint32_t),
                                hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(
        hipMemcpyAsync(d.slots, h.slots.data(), h.slots.size() * sizeof(BatchSlot), hipMemcpyHostToDevice, st));
    const int cus = ctx->compute_units() > 0 ? ctx->compute_units() : 256;
    if (enqueue_plan(h, d, st, cus, t.batch_blocks_per_cu, ctx->profiling ? ctx->prof_start : nullptr,
                     ctx->profiling ? ctx->prof_stop : nullptr) != 0)
        return -1;
    PHIP_HIP_TRY(hipMemcpyAsync(h_res, d.residues, h.slots.size() * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    PHIP_HIP_TRY(hipStreamSynchronize(st));
    if (ctx->profiling) {
        float ms = 0;
        PHIP_HIP_TRY(hipEventElapsedTime(&ms, ctx->prof_start, ctx->prof_stop));
        ctx->profiled_ms += ms;
        ctx->profiled_launches += h.classes.size();
    }
    memcpy(out, h_res, h.slots.size() * sizeof(int64_t));
    return 0;
}

namespace {

// The plan inputs of a batch description (phip_batch_create's arguments).
int parse_batch(const int64_t* entries, const int32_t* orders, int count, const int64_t* moduli,
                const int32_t* moduli_counts, std::vector<PlanInput>& in) {
    if (count < 0 || (count > 0 && (!entries || !orders || !moduli || !moduli_counts))) {
        set_error("invalid batch description");
        return -1;
    }
    in.clear();
    uint64_t entry_off = 0;
    size_t mod_off = 0;
    for (int m = 0; m < count; ++m) {
        const int n = orders[m];
        if (n < kBatchMinOrder || n > kBatchMaxOrder) {
            set_error("batch orders must lie in [5, 46]; use the single-matrix entry point otherwise");
            return -1;
        }
        if (moduli_counts[m] < 0) {
            set_error("negative modulus count");
            return -1;
        }
        for (int k = 0; k < moduli_counts[m]; ++k) {
            const int64_t q = moduli[mod_off + k];
            if (q < 3 || q >= (INT64_C(1) << 31) || !(q & 1)) {
                set_error("batch moduli must be odd and lie in [3, 2^31)");
                return -1;
            }
            in.push_back({entries + entry_off, n, (uint32_t)q, m});
        }
        entry_off += (uint64_t)n * n;
        mod_off += moduli_counts[m];
    }
    return 0;
}

int build_plan(BatchPlan& plan, const int64_t* entries, const int32_t* orders, int count, const int64_t* moduli,
               const int32_t* moduli_counts) {
    std::vector<PlanInput> in;
    if (parse_batch(entries, orders, count, moduli, moduli_counts, in) != 0) return -1;
    if (build_plan_host(in, tuning().batch_items_log2, 4, plan.host) != 0) return -1;
    return plan.upload();
}

}  // namespace

}  // namespace phip

using phip::BatchPlan;

extern "C" {

int phip_batch_create(const int64_t* entries, const int32_t* orders, int count, const int64_t* moduli,
                      const int32_t* moduli_counts, void** out) {
    if (!out) {
        phip::set_error("output handle pointer is null");
        return -1;
    }
    *out = nullptr;
    phip::DeviceContext* ctx = phip::DeviceContext::current();
    if (!ctx) return -1;
    std::unique_ptr<BatchPlan> plan(new BatchPlan());
    plan->device = ctx->device();
    plan->cus = ctx->compute_units() > 0 ? ctx->compute_units() : 256;
    plan->blocks_per_cu = phip::tuning().batch_blocks_per_cu;
    PHIP_HIP_TRY(hipStreamCreateWithFlags(&plan->stream, hipStreamNonBlocking));
    PHIP_HIP_TRY(hipEventCreate(&plan->ev0));
    PHIP_HIP_TRY(hipEventCreate(&plan->ev1));
    if (phip::build_plan(*plan, entries, orders, count, moduli, moduli_counts) != 0) return -1;
    *out = plan.release();
    return 0;
}

int phip_batch_run(void* handle, int64_t* residues_out) {
    if (!handle) {
        phip::set_error("null batch handle");
        return -1;
    }
    auto* plan = static_cast<BatchPlan*>(handle);
    if (plan->host.slots.empty()) return 0;
    return plan->run(residues_out);
}

int phip_batch_launch(void* handle) {
    if (!handle) {
        phip::set_error("null batch handle");
        return -1;
    }
    auto* plan = static_cast<BatchPlan*>(handle);
    if (plan->host.slots.empty()) return 0;
    std::lock_guard<std::mutex> lock(plan->mu);
    if (plan->launched) {
        phip::set_error("batch plan is already running");
        return -1;
    }
    return plan->launch_locked();
}

int phip_batch_query(void* handle, int64_t* residues_out, int block) {
    if (!handle) {
        phip::set_error("null batch handle");
        return -1;
    }
    auto* plan = static_cast<BatchPlan*>(handle);
    if (plan->host.slots.empty()) return 1;
    std::lock_guard<std::mutex> lock(plan->mu);
    return plan->finish_locked(residues_out, block != 0);
}

int phip_batch_set_option(void* handle, int option, int64_t value) {
    if (!handle) {
        phip::set_error("null batch handle");
        return -1;
    }
    auto* plan = static_cast<BatchPlan*>(handle);
    std::lock_guard<std::mutex> lock(plan->mu);
    if (plan->launched) {
        phip::set_error("cannot change options while the plan is running");
        return -1;
    }
    // The captured graph holds the old launch shape and device pointers
    // (re-sizing frees and reallocates the work list), so drop it before
    // anything changes: the next launch captures afresh, also after an
    // option change that fails half-way.
    plan->drop_graph();
    switch (option) {
        case PHIP_BATCH_BLOCKS_PER_CU:
            if (value < 1 || value > 64) {
                phip::set_error("blocks per CU must lie in [1, 64]");
                return -1;
            }
            plan->blocks_per_cu = (int)value;
            break;
        case PHIP_BATCH_ITEMS_LOG2:
            if (value < 0 || value > 30) {
                phip::set_error("work items (log2) must lie in [0, 30]");
                return -1;
            }
            if (plan->resize_items((int)value) != 0) return -1;
            break;
        case PHIP_BATCH_USE_GRAPH:
            plan->use_graph = value != 0;
            break;
        default:
            phip::set_error("unknown batch option");
            return -1;
    }
    return 0;
}

int phip_batch_stats(void* handle, double* out, int len) {
    if (!handle || !out) {
        phip::set_error("null argument");
        return -1;
    }
    auto* plan = static_cast<BatchPlan*>(handle);
    const phip::PlanHost& h = plan->host;
    int cbits_min = h.jobs.empty() ? -1 : 64, cbits_max = -1;
    // Most workgroups per CU any of the plan's class kernels keeps resident:
    // each class's grid is capped at its own count, so a larger setting
    // changes nothing.
    int resident = -1;
    for (const phip::BatchClass& cl : h.classes) resident = std::max(resident, phip::class_resident(cl));
    for (const phip::BatchJob& jb : h.jobs) {
        cbits_min = std::min(cbits_min, jb.cbits);
        cbits_max = std::max(cbits_max, jb.cbits);
    }
    const double v[] = {(double)plan->last_ms,
                        h.total_states,
                        h.total_mulmods,
                        (double)h.jobs.size(),
                        (double)h.items.size(),
                        (double)h.classes.size(),
                        (double)(plan->cus * plan->blocks_per_cu * 4),
                        h.classes.empty() ? -1.0 : (double)h.classes[0].group,
                        (double)h.slots.size(),
                        h.classes.empty() ? -1.0 : (double)h.classes[0].set,
                        (double)cbits_min,
                        (double)cbits_max,
                        (double)h.items_log2,
                        (double)plan->blocks_per_cu,
                        h.classes.empty() ? -1.0 : (double)h.classes[0].variant,
                        (double)resident,
                        h.jobs.empty() ? -1.0 : (double)h.jobs[0].test_words,
                        h.jobs.empty() ? -1.0 : (double)h.jobs[0].block_rows};
    const int m = std::min<int>(len, (int)(sizeof(v) / sizeof(v[0])));
    for (int i = 0; i < m; ++i) out[i] = v[i];
    return m;
}

int phip_batch_plan_classes(const int64_t* entries, const int32_t* orders, int count, const int64_t* moduli,
                            const int32_t* moduli_counts, int32_t* out, int len) {
    if (!out && len > 0) {
        phip::set_error("null argument");
        return -1;
    }
    std::vector<phip::PlanInput> in;
    if (phip::parse_batch(entries, orders, count, moduli, moduli_counts, in) != 0) return -1;
    phip::PlanHost h;
    if (phip::build_plan_host(in, phip::tuning().batch_items_log2, 4, h) != 0) return -1;
    const int m = std::min<int>(len, (int)h.slots.size());
    for (int i = 0; i < m; ++i) {
        const phip::BatchJob& jb = h.jobs[h.slots[i].job];
        out[i] = (jb.group * 32 + jb.set) * 8 + jb.variant;
    }
    return (int)h.slots.size();
}

int64_t phip_block_tables(const int64_t* a, int n, int mod3, int32_t* out, int64_t cap, int32_t* meta) {
    if (!a || !meta || (!out && cap > 0)) {
        phip::set_error("null argument");
        return -1;
    }
    if (n < phip::kBlockMinOrder || n > phip::kBatchMaxOrder) {
        phip::set_error("block tables take orders 8..46");
        return -1;
    }
    for (size_t e = 0; e < (size_t)n * n; ++e)
        if (a[e] != 0 && a[e] != 1 && !(mod3 && a[e] == -1)) {
            phip::set_error(mod3 ? "modulo-3 block tables take entries -1, 0 and 1" : "block tables take 0/1 matrices");
            return -1;
        }
    std::vector<int> order;
    int group = 0;
    const int even = phip::block_plan(a, n, order, group);
    if (even < 0) {
        phip::set_error("no row-sum group fits the row sums");
        return -1;
    }
    const int rows = mod3 ? phip::block_test_rows(n, n) : phip::block_test_rows(n, even);
    const size_t words = mod3 ? phip::block_table_words_mod3(n, rows) : phip::block_table_words(n, rows);
    meta[0] = rows;
    meta[1] = even;
    meta[2] = group;
    meta[3] = n;
    if ((int64_t)words <= cap) {
        if (mod3) {
            phip::build_block_tables_mod3(a, n, order, rows, out);
        } else {
            phip::build_block_tables(a, n, order, even, rows, out);
        }
    }
    return (int64_t)words;
}

int phip_batch_destroy(void* handle) {
    delete static_cast<BatchPlan*>(handle);
    return 0;
}

}  // extern "C"
