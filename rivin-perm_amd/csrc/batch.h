// Shared declarations of the persistent Glynn engine (batch.hip,
// batch_plan.hip).  Every 31-bit residue of the library runs through it: a
// batch plan keeps its tables resident on the device, the single-matrix
// entry points build a one-matrix plan in the device context's buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace phip {

class DeviceContext;

constexpr int kBatchMinOrder = 5;
constexpr int kBatchMaxOrder = 46;
// Most primes one job evaluates together.
constexpr int kMaxSet = 8;
// Per-item partial sums: one 64-bit slot per prime of the job's set, or per
// limb column of an exact job (at most kExactMaxColumns).
constexpr int kItemSlots = 12;

// Kernel variants of a (group, set) class.
constexpr int kVariantPlain = 0;
constexpr int kVariantPaired = 1;     // consecutive states paired (primes below kPairedPrimeLimit)
constexpr int kVariantSkipZeros = 2;  // zero terms skipped through a per-wavefront queue
constexpr int kVariantExact = 3;      // exact accumulation: one walk for any number of primes
constexpr int kVariantExactCompact = 4;  // the same with the product tree sized for kCompactValueBits-bit values
constexpr int kVariantSkipExact = 5;  // the block zero-skipping walk with exact accumulation (kExactColumns columns)
constexpr int kVariantExactSigned = 6;  // exact accumulation of signed group products in signed limbs (8 columns)
// Value-bits parameter of the exact sink that selects signed limbs.
constexpr int kSignedLimbs = 0;
// Bound of a four-group product for the signed-limb class (top limbs of
// the four-limb operands below 2^29).
constexpr int kSignedQuadBits = 122;
// The exact variant's classes: a class's set size is its number of limb
// columns L.  L = 8 takes tables of at most eight groups (full products);
// L = 10 and 12 take up to sixteen groups when the product of the table's
// row bounds is below 2^(31 L) (products kept modulo 2^(31 L)).
constexpr int kExactColumns = 8;
constexpr int kExactMaxGroups = 8;
constexpr int kExactWideMaxGroups = 16;
constexpr int kExactMaxColumns = 12;
// The compact exact classes (groups of 2 rows, 10 limb columns): a product
// of c >= 2 group values keeps ceil(22 c / 31) limbs (LimbProduct with
// V = 22), for tables whose partial bound products fit (the products of two
// row sums of entries up to about 100 at orders 24..28 are near 2^21).
constexpr int kCompactValueBits = 22;
constexpr bool is_exact_variant(int variant) {
    return variant == kVariantExact || variant == kVariantExactCompact || variant == kVariantSkipExact ||
           variant == kVariantExactSigned;
}
// Variants that walk 0/1 tables on the zero-skipping kernels (byte rows,
// block tables, per-wavefront LDS regions).
constexpr bool is_skip_variant(int variant) { return variant == kVariantSkipZeros || variant == kVariantSkipExact; }

// One job: one table (a matrix under one row order) and a set of K odd
// moduli below 2^31 that share it (K = 1 for residue rows).  56 bytes,
// read with scalar loads.
struct BatchJob {
    uint64_t data_off;   // int32 offset of the job's tables (n * glynn_stride words; n * byte_stride for kVariantSkipZeros)
    uint64_t lanes;      // lanes of this job (2^(n-1-cbits))
    int32_t n;
    int32_t cbits;
    uint32_t prime_off;  // first of the job's K entries in the prime arrays
    uint32_t first_item;
    uint32_t n_items;
    int32_t group;
    int32_t nprimes;     // real primes (<= K); padded entries repeat the last one
    int32_t set;         // K of the job's launch class
    int32_t variant;     // kVariantPlain, kVariantPaired, kVariantSkipZeros, kVariantExact, kVariantExactCompact
                         // kVariantSkipExact or kVariantExactSigned
    int32_t test_words;  // kVariantSkipZeros: leading table words the zero test reads (skip_test_words)
    int32_t block_rows;  // kVariantSkipZeros, kVariantSkipExact: rows the 32-state block test reads
                         // (block_test_rows: a multiple of 8 or every row word; the rows of even degree lead
                         // the table), or -1 for the per-state test
    int32_t block_mod3;  // kVariantSkipZeros with block_rows >= 0: 1 for the modulo-3 block walk (row sums that
                         // are multiples of 3 zero a term; build_block_tables_mod3), else 0
};

// One requested residue: prime slot `k` of job `job`.
struct BatchSlot {
    uint32_t job;
    uint32_t k;
    uint32_t p;
    uint32_t scale;  // R^groups * 2^-(n-1) mod p (exact jobs: 2^-(n-1) mod p)
};

// One persistent launch: the jobs of one (row-sum group, set size,
// variant) class.
struct BatchClass {
    int group;
    int set;
    int variant;
    uint32_t first_item;
    uint32_t n_items;
    int max_n;  // largest order of the class's jobs (sizes the zero-skip queues)
    int max_block_words;  // largest block-walk LDS region of the class's jobs (block_region_words), or 0
};

// Rows of one wavefront's zero-skip queue for tables of order n: the byte
// row words and the sign (ZeroSkipSink), kZeroQueueEntries words each.
constexpr int kZeroQueueEntries = 128;
constexpr int skip_queue_rows(int n) { return (n + 3) / 4 + 1; }
// Word counts the zero-skipping walk's test may read for tables of order n:
// every word, or the leading half of them when every row of even degree
// lies there (a row of odd degree never sums to zero).
// The half-width test is compiled for orders from kSkipHalfMinOrder on
// (below it a residue takes microseconds either way).
constexpr int kSkipHalfMinOrder = 20;
constexpr int skip_full_words(int n) { return (n + 3) / 4; }
constexpr int skip_half_words(int n) { return n < kSkipHalfMinOrder ? (n + 3) / 4 : ((n + 3) / 4 + 1) / 2; }
// The block form of the zero-skipping walk (glynn_chunk_set_blocks): the
// kBlockInnerBits fastest Gray columns are enumerated per outer state, and
// one 32-bit LDS lookup per row marks the inner states that zero the row.
// Its per-wavefront LDS region (words): the rows' mask tables (32 words a
// row, plus 32), the inner row-sum tables (12 words per row word: 4 for
// the first two inner columns, 8 for the other three), 64 words of owner
// flags and kBlockRing entries of (row words, survivor mask, first state
// index).
constexpr int kBlockInnerBits = 5;
constexpr int kBlockRing = 128;
constexpr int kBlockMinChunkBits = kBlockInnerBits + 1;
// Smallest order planned on the block walk (smaller tables take the
// per-state test).
constexpr int kBlockMinOrder = 8;
constexpr int block_region_words(int n, int rows) {
    return (rows + 1) * 32 + 12 * ((n + 3) / 4) + 64 + ((n + 3) / 4 + 2) * kBlockRing;
}
// The modulo-3 block walk's region: 64-word mask tables per tested row and
// the inner tables; its rare nonzero states are taken in place, with no
// ring.
constexpr int block_region_words_mod3(int n, int rows) { return rows * 64 + 12 * ((n + 3) / 4); }
// Dynamic LDS of one workgroup (4 wavefronts) of a class's kernel: the
// zero-skip queues or block-walk regions (only block regions for the exact
// form, which runs block tables only), nothing for the other variants.
inline unsigned class_lds_bytes(const BatchClass& cl) {
    if (cl.variant == kVariantSkipExact) return 4u * 4u * (unsigned)cl.max_block_words;
    if (cl.variant != kVariantSkipZeros) return 0u;
    const int words = skip_queue_rows(cl.max_n) * kZeroQueueEntries;
    return 4u * 4u * (unsigned)(words > cl.max_block_words ? words : cl.max_block_words);
}

// One requested residue before planning: a row-major n x n matrix and an odd
// modulus in [3, 2^31).  Inputs with equal `key` may share a job when their
// tables are identical.
struct PlanInput {
    const int64_t* a;
    int n;
    uint32_t q;
    int key;
    // The modulo-3 block walk: q == 3, entries in {-1, 0, 1}, order
    // kBlockMinOrder..kBatchMaxOrder.
    bool mod3 = false;
};

// Host image of a plan.
struct PlanHost {
    std::vector<BatchJob> jobs;
    std::vector<uint2> items;  // (job, wavefront index), grouped by class
    std::vector<int32_t> blob;
    std::vector<int32_t> set_p;
    std::vector<uint32_t> set_pinv;
    std::vector<BatchSlot> slots;  // input order
    std::vector<BatchClass> classes;
    double total_states = 0;   // Gray states over all residues
    double total_mulmods = 0;  // n 2^(n-1) per residue
    int items_log2 = 0;        // work-item target the chunks were sized for
    int min_cbits = 0;
};

// Device image of a plan (pointers into device memory).
struct PlanDevice {
    BatchJob* jobs;
    uint2* items;
    int32_t* blob;
    int32_t* set_p;
    uint32_t* set_pinv;
    BatchSlot* slots;
    uint64_t* partials;  // kItemSlots per item
    uint32_t* counters;  // one per class
    int64_t* residues;   // one per slot
};

// Validates the inputs and builds the plan: tables per distinct (matrix,
// row order, representation), prime sets, cost-sorted items of one
// wavefront each, about 2^items_log2 items in all.  Returns 0 or -1.
int build_plan_host(const std::vector<PlanInput>& in, int items_log2, int min_cbits, PlanHost& out);

// Re-sizes the chunks of an existing plan's jobs for about 2^items_log2
// work items (chunk bits, lanes, items and classes; the tables, prime sets
// and slots stay).  Returns 0 or -1.
int assign_items(PlanHost& out, int items_log2, int min_cbits);

// Enqueues one run: counter reset, one persistent launch per class (between
// ev0 and ev1 when given; cus x min(blocks_per_cu, what the class's kernel
// keeps resident) workgroups), finalisation into d.residues.  Returns 0 or -1.
int enqueue_plan(const PlanHost& h, const PlanDevice& d, hipStream_t stream, int cus, int blocks_per_cu,
                 hipEvent_t ev0, hipEvent_t ev1);

// Workgroups per CU a class's kernel keeps resident with its dynamic LDS,
// or -1.
int class_resident(const BatchClass& cl);

// Builds and runs a plan for `in` in the device context's stream and
// buffers (the single-matrix entry points); out[i] receives the residue of
// in[i].  Chunk sizes follow the context tuning.  Returns 0 or -1.
int run_plan_in_context(DeviceContext* ctx, const std::vector<PlanInput>& in, int64_t* out);

// Launches a class's persistent kernel; lds_bytes of dynamic LDS per
// workgroup (class_lds_bytes).
using BatchLaunchFn = hipError_t (*)(dim3 grid, unsigned lds_bytes, hipStream_t stream, const BatchJob* jobs,
                                     const uint2* items, uint32_t n_items, uint32_t* counter, const int32_t* blob,
                                     const int32_t* set_p, const uint32_t* set_pinv, uint64_t* item_partials);

#define PHIP_DECLARE_BATCH_LAUNCH(G, K)                                                                    \
    hipError_t launch_batch_g##G##_k##K(dim3, unsigned, hipStream_t, const BatchJob*, const uint2*, uint32_t, uint32_t*, \
                                        const int32_t*, const int32_t*, const uint32_t*, uint64_t*);
PHIP_DECLARE_BATCH_LAUNCH(0, 1)
PHIP_DECLARE_BATCH_LAUNCH(2, 1)
PHIP_DECLARE_BATCH_LAUNCH(2, 2)
PHIP_DECLARE_BATCH_LAUNCH(2, 4)
PHIP_DECLARE_BATCH_LAUNCH(2, 8)
PHIP_DECLARE_BATCH_LAUNCH(3, 1)
PHIP_DECLARE_BATCH_LAUNCH(3, 2)
PHIP_DECLARE_BATCH_LAUNCH(3, 4)
PHIP_DECLARE_BATCH_LAUNCH(3, 8)
PHIP_DECLARE_BATCH_LAUNCH(4, 1)
PHIP_DECLARE_BATCH_LAUNCH(4, 2)
PHIP_DECLARE_BATCH_LAUNCH(4, 4)
PHIP_DECLARE_BATCH_LAUNCH(4, 8)
PHIP_DECLARE_BATCH_LAUNCH(6, 1)
PHIP_DECLARE_BATCH_LAUNCH(6, 2)
PHIP_DECLARE_BATCH_LAUNCH(6, 4)
PHIP_DECLARE_BATCH_LAUNCH(6, 8)
#undef PHIP_DECLARE_BATCH_LAUNCH
// Paired-state variants (every prime of the set below kPairedPrimeLimit).
#define PHIP_DECLARE_BATCH_LAUNCH_PAIRED(G, K)                                                                   \
    hipError_t launch_batch_g##G##_k##K##_paired(dim3, unsigned, hipStream_t, const BatchJob*, const uint2*, uint32_t,   \
                                               uint32_t*, const int32_t*, const int32_t*, const uint32_t*,   \
                                               uint64_t*);
PHIP_DECLARE_BATCH_LAUNCH_PAIRED(2, 4)
PHIP_DECLARE_BATCH_LAUNCH_PAIRED(2, 8)
PHIP_DECLARE_BATCH_LAUNCH_PAIRED(3, 4)
PHIP_DECLARE_BATCH_LAUNCH_PAIRED(3, 8)
PHIP_DECLARE_BATCH_LAUNCH_PAIRED(4, 4)
PHIP_DECLARE_BATCH_LAUNCH_PAIRED(4, 8)
PHIP_DECLARE_BATCH_LAUNCH_PAIRED(6, 4)
PHIP_DECLARE_BATCH_LAUNCH_PAIRED(6, 8)
#undef PHIP_DECLARE_BATCH_LAUNCH_PAIRED
// Zero-skipping variants (0/1 tables whose states are mostly zero terms).
#define PHIP_DECLARE_BATCH_LAUNCH_SKIP(G, K)                                                                   \
    hipError_t launch_batch_g##G##_k##K##_skip(dim3, unsigned, hipStream_t, const BatchJob*, const uint2*, uint32_t,   \
                                             uint32_t*, const int32_t*, const int32_t*, const uint32_t*,   \
                                             uint64_t*);
PHIP_DECLARE_BATCH_LAUNCH_SKIP(4, 2)
PHIP_DECLARE_BATCH_LAUNCH_SKIP(4, 4)
PHIP_DECLARE_BATCH_LAUNCH_SKIP(4, 8)
PHIP_DECLARE_BATCH_LAUNCH_SKIP(6, 2)
PHIP_DECLARE_BATCH_LAUNCH_SKIP(6, 4)
PHIP_DECLARE_BATCH_LAUNCH_SKIP(6, 8)
#undef PHIP_DECLARE_BATCH_LAUNCH_SKIP
// The block zero-skipping walk with exact accumulation (groups of 6, 8 limb
// columns).
hipError_t launch_batch_g6_k8_skipexact(dim3, unsigned, hipStream_t, const BatchJob*, const uint2*, uint32_t,
                                        uint32_t*, const int32_t*, const int32_t*, const uint32_t*, uint64_t*);
// Exact-accumulation variants (any number of primes; L = K limb columns).
#define PHIP_DECLARE_BATCH_LAUNCH_EXACT(G, K)                                                                 \
    hipError_t launch_batch_g##G##_k##K##_exact(dim3, unsigned, hipStream_t, const BatchJob*, const uint2*, uint32_t,  \
                                              uint32_t*, const int32_t*, const int32_t*, const uint32_t*,  \
                                              uint64_t*);
PHIP_DECLARE_BATCH_LAUNCH_EXACT(2, 8)
PHIP_DECLARE_BATCH_LAUNCH_EXACT(3, 8)
PHIP_DECLARE_BATCH_LAUNCH_EXACT(4, 8)
PHIP_DECLARE_BATCH_LAUNCH_EXACT(6, 8)
PHIP_DECLARE_BATCH_LAUNCH_EXACT(2, 10)
PHIP_DECLARE_BATCH_LAUNCH_EXACT(2, 12)
PHIP_DECLARE_BATCH_LAUNCH_EXACT(3, 10)
PHIP_DECLARE_BATCH_LAUNCH_EXACT(3, 12)
PHIP_DECLARE_BATCH_LAUNCH_EXACT(4, 10)
PHIP_DECLARE_BATCH_LAUNCH_EXACT(4, 12)
#undef PHIP_DECLARE_BATCH_LAUNCH_EXACT
// Signed-limb exact classes (groups of 3 or 4 rows, 8 limb columns).
hipError_t launch_batch_g3_k8_signed(dim3, unsigned, hipStream_t, const BatchJob*, const uint2*, uint32_t, uint32_t*,
                                     const int32_t*, const int32_t*, const uint32_t*, uint64_t*);
hipError_t launch_batch_g4_k8_signed(dim3, unsigned, hipStream_t, const BatchJob*, const uint2*, uint32_t, uint32_t*,
                                     const int32_t*, const int32_t*, const uint32_t*, uint64_t*);
// The compact exact class (groups of 2 rows, 10 limb columns).
hipError_t launch_batch_g2_k10_compact(dim3, unsigned, hipStream_t, const BatchJob*, const uint2*, uint32_t, uint32_t*,
                                       const int32_t*, const int32_t*, const uint32_t*, uint64_t*);

// Workgroups per CU each class kernel can keep resident with lds_bytes of
// dynamic LDS per workgroup (batch.hip), or -1.
#define PHIP_DECLARE_BATCH_OCC(NAME) int NAME(unsigned lds_bytes);
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g0_k1)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g2_k1) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g2_k2)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g2_k4) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g2_k8)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g3_k1) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g3_k2)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g3_k4) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g3_k8)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g4_k1) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g4_k2)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g4_k4) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g4_k8)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g6_k1) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g6_k2)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g6_k4) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g6_k8)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g2_k4_paired) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g2_k8_paired)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g3_k4_paired) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g3_k8_paired)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g4_k4_paired) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g4_k8_paired)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g6_k4_paired) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g6_k8_paired)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g4_k2_skip) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g4_k4_skip)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g4_k8_skip) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g6_k2_skip)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g6_k4_skip) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g6_k8_skip)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g2_k8_exact) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g3_k8_exact)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g4_k8_exact) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g6_k8_exact)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g2_k10_exact) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g2_k12_exact)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g3_k10_exact) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g3_k12_exact)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g4_k10_exact) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g4_k12_exact)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g2_k10_compact)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g6_k8_skipexact)
PHIP_DECLARE_BATCH_OCC(occupancy_batch_g3_k8_signed) PHIP_DECLARE_BATCH_OCC(occupancy_batch_g4_k8_signed)
#undef PHIP_DECLARE_BATCH_OCC

}  // namespace phip
