// Internal residue engine shared by the C-ABI entry points.
#pragma once
#include <stdint.h>

#include <vector>

#include "device_context.h"

namespace phip {

// Tunables (set through phip_set_tuning): target number of lanes per
// single-matrix launch; minimum chunk bits; batch item count and workgroups
// per CU; the smallest prime Schur order and the task depth of the necklace
// kernel; the smallest row-register tiers of the wide and affine Glynn
// kernels; the largest estimated share of nonzero terms (permille) at which
// a 0/1 table runs on the zero-skipping kernels (0 turns them off).
struct Tuning {
    int lanes_log2 = 20;
    int min_cbits = 4;
    int batch_items_log2 = 15;
    int batch_blocks_per_cu = 4;  // capped per class at what its kernel keeps resident
    int necklace_min_order = 23;  // smallest prime Schur order sent to the necklace kernel
    int necklace_depth = 28;      // largest FKM prefix depth of the necklace tasks
    int wide_min_tier = 8;        // smallest register tier (8..64, multiple of 4) of the wide Glynn kernels
    int affine_min_tier = 8;      // smallest register tier (8..40, multiple of 4) of the affine Glynn kernels
    int skip_zeros_permille = 250;
    int exact_min_primes = 6;     // fewest primes of a table of <= 8 groups that take the exact variant (0 off)
    int exact_wide_min_primes = 6;  // the same for 9..16 groups (row-bound product below 2^372)
    int exact_compact = 1;          // groups of 2 rows may take the compact exact classes (0 off)
    int skip_blocks = 1;            // zero-skipping tables test 32 Gray states per row lookup (0: per state)
    int mod3_blocks = 1;            // permanents modulo 3 of orders 8..46 on the modulo-3 block walk (0: bit-sliced)
    int exact_signed = 1;           // groups of 3 or 4 rows may take the signed-limb exact classes (0 off)
    int skip_exact_min_primes = 5;  // fewest primes of a 0/1 block table of <= 8 groups on the exact block walk (0 off)
};
Tuning& tuning();

// perm(A_i) mod moduli[i] for i < k, where A_i = A + i * stride (row-major
// n x n int64).  Any 0 < m < 2^63.  Returns 0 or -1 (error set).
int compute_residues(DeviceContext* ctx, const int64_t* A, int64_t stride, int n,
                     const int64_t* moduli, int k, int64_t* out);

// Deterministic Miller-Rabin for 64-bit values.
bool is_prime_u64(uint64_t v);

}  // namespace phip
