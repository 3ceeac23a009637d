// Host construction of the per-(matrix, modulus) tables read by glynn_chunk.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "batch.h"
#include "launchers.h"
#include "phip_arith.h"

namespace phip {

// The representative of v modulo an odd q in [-(q-1)/2, (q-1)/2]: v itself
// when it already lies there (the common case, no division).
inline int64_t centred_mod(int64_t v, int64_t q) {
    const int64_t h = (q - 1) / 2;
    if (v >= -h && v <= h) return v;
    const int64_t r = host_floor_mod(v, q);
    return r > h ? r - q : r;
}

// Row bounds S_i = sum_j |centred a_ij mod q|.
inline std::vector<uint64_t> centred_row_bounds(const int64_t* M, int n, int64_t q) {
    std::vector<uint64_t> S(n, 0);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
            const int64_t v = centred_mod(M[(size_t)r * n + c], q);
            S[r] += (uint64_t)(v < 0 ? -v : v);
        }
    return S;
}

// Saturating product of row bounds (values above 2^62 collapse to 2^62).
inline uint64_t bound_product(const uint64_t* b, int c) {
    unsigned __int128 p = 1;
    for (int i = 0; i < c; ++i) {
        p *= b[i];
        if (p > ((unsigned __int128)1 << 62)) return (uint64_t)1 << 62;
    }
    return (uint64_t)p;
}

// Whether `c` row bounds in this order satisfy the operand limits of the
// kernel's product of a group of c rows (glynn31.h, PackedRows::run_product):
// every partial product that feeds a 24-bit multiply below 2^23, the group
// product below 2^31.
inline bool group_fits(const uint64_t* b, int c) {
    const uint64_t k23 = 1ull << 23, k31 = 1ull << 31;
    if (bound_product(b, c) >= k31) return false;
    switch (c) {
        case 1:
        case 2: return true;
        case 3: return bound_product(b, 2) < k23;
        case 4: return bound_product(b, 2) < k23 && bound_product(b + 2, 2) < k23;
        case 5: return bound_product(b, 3) < k23 && bound_product(b + 3, 2) < k23;
        case 6: return bound_product(b, 3) < k23 && bound_product(b + 3, 3) < k23;
        default: return false;
    }
}

// Row order and group size for the packed kernels: rows are dealt largest
// bound first into the group with the smallest bound product so far (all
// groups hold G rows except the last), then each group is arranged so its
// partial products fit.  Row order does not change the permanent.  Returns
// the largest feasible G in {6, 4, 3, 2}, or 0 (wide residue rows).
inline int plan_rows(const std::vector<uint64_t>& S, std::vector<int>& order) {
    const int n = (int)S.size();
    order.resize(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    for (uint64_t s : S)
        if (s >= (1ull << 15)) return 0;
    std::vector<int> by_size(order);
    std::stable_sort(by_size.begin(), by_size.end(), [&](int a, int b) { return S[a] > S[b]; });
    for (int G : {6, 4, 3, 2}) {
        const int ng = (n + G - 1) / G;
        std::vector<std::vector<int>> groups(ng);
        std::vector<double> logp(ng, 0.0);
        for (int r : by_size) {
            int best = -1;
            for (int g = 0; g < ng; ++g) {
                const int cap = (g == ng - 1) ? n - (ng - 1) * G : G;
                if ((int)groups[g].size() >= cap) continue;
                if (best < 0 || logp[g] < logp[best]) best = g;
            }
            groups[best].push_back(r);
            logp[best] += std::log2((double)S[r] + 1.0);
        }
        bool ok = true;
        std::vector<int> out;
        for (auto& grp : groups) {
            std::sort(grp.begin(), grp.end());
            // The group product's bound does not depend on the arrangement:
            // a group above it fails every one of them.
            uint64_t all[6];
            for (size_t i = 0; i < grp.size(); ++i) all[i] = S[grp[i]];
            if (bound_product(all, (int)grp.size()) >= (1ull << 31)) {
                ok = false;
                break;
            }
            bool found = false;
            do {
                uint64_t b[6];
                for (size_t i = 0; i < grp.size(); ++i) b[i] = S[grp[i]];
                if (group_fits(b, (int)grp.size())) {
                    found = true;
                    break;
                }
            } while (std::next_permutation(grp.begin(), grp.end()));
            if (!found) {
                ok = false;
                break;
            }
            out.insert(out.end(), grp.begin(), grp.end());
        }
        if (ok) {
            order = out;
            return G;
        }
    }
    return 0;
}

// Writes n * glynn_stride(n, group) int32 words: the row sums of state 0
// (every column sign +1), then the flip deltas 2a_ij of the n-1 Gray columns,
// with kernel row i taken from matrix row order[i] (order may be empty for
// the identity).  group 0: residues in [0, q) one per word.  group > 0: exact
// centred values as 16-bit two's-complement halves, row 2k in the low half of
// word k.
inline void build_glynn_tables(const int64_t* M, int n, uint32_t q, int group, const std::vector<int>& order,
                               int32_t* out) {
    const int W = glynn_stride(n, group);
    std::vector<int64_t> v((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const int src = order.empty() ? i : order[i];
            const int64_t a = M[(size_t)src * n + j];
            if (group == 0) {
                v[(size_t)i * n + j] = (a >= 0 && a < (int64_t)q) ? a : host_floor_mod(a, (int64_t)q);
            } else {
                v[(size_t)i * n + j] = centred_mod(a, (int64_t)q);
            }
        }
    auto put = [&](int col, int row, int64_t value) {
        int32_t* c = out + (size_t)col * W;
        if (group == 0) {
            c[row] = (int32_t)value;
        } else {
            uint32_t w = (uint32_t)c[row / 2];
            const uint32_t h = (uint32_t)(uint16_t)(int16_t)value;
            w = (row & 1) ? ((w & 0xffffu) | (h << 16)) : ((w & 0xffff0000u) | h);
            c[row / 2] = (int32_t)w;
        }
    };
    for (size_t i = 0; i < (size_t)n * W; ++i) out[i] = 0;
    for (int i = 0; i < n; ++i) {
        int64_t row = 0;
        for (int j = 0; j < n; ++j) row = group == 0 ? (row + v[(size_t)i * n + j]) % (int64_t)q : row + v[(size_t)i * n + j];
        put(0, i, row);
        for (int j = 0; j < n - 1; ++j) {
            const int64_t a = v[(size_t)i * n + j];
            put(1 + j, i, group == 0 ? (2 * a) % (int64_t)q : 2 * a);
        }
    }
    // The padding half of an odd packed order holds 1 in every state (its
    // flip deltas are 0), so a halfwise minimum over the row words sees a
    // zero only for a zero row sum.  No product reads it.
    if (group > 0 && (n & 1)) put(0, n, 1);
}

// Tables of the zero-skipping walk (glynn31.h, ByteRows) for a matrix with
// entries in {0, 1}: n * byte_stride(n) int32 words, the biased row sums
// 64 + sum_j a_ij of state 0 (every column sign +1), then the flip deltas
// 2 a_ij of the n-1 Gray columns, one byte per row (kernel row i from
// matrix row order[i], in byte i % 4 of word i / 4).  Padding rows hold 65
// (a row sum of 1, never zero) with zero deltas.  Every row degree must be
// at most 63, so that each row sum stays in [-63, 63].
inline void build_byte_tables(const int64_t* M, int n, const std::vector<int>& order, int32_t* out) {
    const int W = byte_stride(n);
    std::vector<uint32_t> words((size_t)n * W, 0u);
    auto put = [&](int col, int row, uint32_t byte) { words[(size_t)col * W + row / 4] |= byte << (8 * (row % 4)); };
    for (int i = 0; i < 4 * W; ++i) {
        if (i >= n) {
            put(0, i, 65u);
            continue;
        }
        const int src = order.empty() ? i : order[i];
        uint32_t degree = 0;
        for (int j = 0; j < n; ++j) degree += M[(size_t)src * n + j] != 0 ? 1u : 0u;
        put(0, i, 64u + degree);
        for (int j = 0; j < n - 1; ++j)
            if (M[(size_t)src * n + j] != 0) put(1 + j, i, 2u);
    }
    for (size_t k = 0; k < words.size(); ++k) out[k] = (int32_t)words[k];
}

}  // namespace phip

namespace phip {

// Rows the block walk tests for a table of order n whose `even` leading
// rows have even degree: whole pairs of row words (the kernel tests eight
// rows at a time), at most every row word.
inline int block_test_rows(int n, int even) {
    const int words = byte_stride(n);
    return 4 * std::min(words, 2 * ((even + 7) / 8));
}

// Words of a block-walk table (build_block_tables) of order n whose `rows`
// leading rows are tested (block_test_rows).
inline size_t block_table_words(int n, int rows) {
    const size_t W = (size_t)byte_stride(n);
    return (size_t)n * W + 12 * W + ((size_t)rows + 1) * 32;
}

// Tables of the block zero-skipping walk (glynn31.h,
// glynn_chunk_set_blocks) for a matrix with entries in {0, 1} whose rows,
// in kernel order order[i], put the `even` rows of even degree first.  Row
// sums are kept doubled, as the byte 2 s + beta_i with beta_i = 128 for a
// row of even degree and 126 for one of odd degree, so every byte is a
// multiple of 4 (an aligned LDS word address) and lies in [0, 254] for
// |s| <= 63; padding rows hold 132.  Layout, in int32 words (W =
// byte_stride(n)):
//   [0, n W)          the bytes of state 0 (every column sign +1), then the
//                     flip deltas 4 a_ij of the n-1 Gray columns;
//   then 12 W         the inner tables, for row word k at 12 k: with I the
//                     deltas summed over the set bits y_j of y (the first
//                     kBlockInnerBits Gray columns), split as I = A(y & 3)
//                     + B(y >> 2) over columns 0-1 and 2-4, and C the
//                     per-byte corrections 128 - beta_i, the 4 words
//                     -A(a) / 2 and the 8 words -((B(b) - C) / 2 - 0x40404040)
//                     (modulo 2^32; A and B - C are even and lie in
//                     (-2^31, 2^31)).  Half an outer word plus both has the
//                     bytes s + 128 of the row sums at (outer state, y):
//                     (O - I + C) / 2 has the bytes s + 64 in [1, 127] (O,
//                     I and C have even bytes), so no byte carries into
//                     the next;
//   then (rows+1) 32  the masks: entry e of row i (i < rows), at word
//                     32 i + e, has bit y set when the row's sum at inner
//                     state y is zero for an outer byte 4 e, i.e. when e =
//                     32 + (the row's ones among the inner columns selected
//                     by y), so only e in [32, 38) is ever nonzero.  Rows
//                     of odd degree never sum to zero: their entries are 0.
//                     Rows overlap (a stride of 32 words, not 64): a row
//                     byte of order n <= 46 lies in [36, 220], so e lies in
//                     [9, 55] and a read of row i meets only its own
//                     nonzero words or words no row sets (row i-1's nonzero
//                     words are its e in [0, 6), row i+1's its e in
//                     [64, 70)).
inline void build_block_tables(const int64_t* M, int n, const std::vector<int>& order, int even, int rows,
                               int32_t* out) {
    const int W = byte_stride(n);
    std::vector<uint32_t> words(block_table_words(n, rows), 0u);
    auto put = [&](size_t word_base, int row, uint32_t byte) { words[word_base + row / 4] |= byte << (8 * (row % 4)); };
    auto entry = [&](int i, int j) { return M[(size_t)order[i] * n + j] != 0 ? 1u : 0u; };
    const size_t inner_at = (size_t)n * W, masks_at = inner_at + 12 * (size_t)W;
    std::vector<uint32_t> correction(W, 0u);
    std::vector<uint32_t> low((size_t)W * 4, 0u), high((size_t)W * 8, 0u);  // A(a), B(b) per row word
    for (int i = 0; i < 4 * W; ++i) {
        if (i >= n) {
            put(0, i, 132u);
            continue;
        }
        uint32_t degree = 0;
        for (int j = 0; j < n; ++j) degree += entry(i, j);
        const uint32_t beta = degree % 2 == 0 ? 128u : 126u;
        put(0, i, 2u * degree + beta);
        for (int j = 0; j < n - 1; ++j)
            if (entry(i, j)) put((size_t)(1 + j) * W, i, 4u);
        correction[i / 4] |= (128u - beta) << (8 * (i % 4));
        for (int a = 0; a < 4; ++a)
            low[(size_t)(i / 4) * 4 + a] |= (4u * ((a & 1 ? entry(i, 0) : 0u) + (a & 2 ? entry(i, 1) : 0u)))
                                            << (8 * (i % 4));
        for (int b = 0; b < 8; ++b)
            high[(size_t)(i / 4) * 8 + b] |=
                (4u * ((b & 1 ? entry(i, 2) : 0u) + (b & 2 ? entry(i, 3) : 0u) + (b & 4 ? entry(i, 4) : 0u)))
                << (8 * (i % 4));
        for (int y = 0; y < 32; ++y) {
            uint32_t ones = 0;
            for (int j = 0; j < kBlockInnerBits; ++j)
                if ((y >> j) & 1) ones += entry(i, j);
            if (i < even && i < rows) words[masks_at + (size_t)i * 32 + 32 + ones] |= 1u << y;
        }
    }
    for (int k = 0; k < W; ++k) {
        for (int a = 0; a < 4; ++a)
            words[inner_at + (size_t)k * 12 + a] = (uint32_t)(-((int64_t)low[(size_t)k * 4 + a] / 2));
        for (int b = 0; b < 8; ++b) {
            const int64_t d = (int64_t)high[(size_t)k * 8 + b] - (int64_t)correction[k];
            words[inner_at + (size_t)k * 12 + 4 + b] = (uint32_t)(-(d / 2 - 0x40404040));
        }
    }
    for (size_t k = 0; k < words.size(); ++k) out[k] = (int32_t)words[k];
}


// Words of a modulo-3 block-walk table (build_block_tables_mod3).
inline size_t block_table_words_mod3(int n, int rows) {
    const size_t W = (size_t)byte_stride(n);
    return (size_t)n * W + 12 * W + (size_t)rows * 64;
}

// Tables of the modulo-3 block walk for a matrix with entries in {-1, 0, 1}
// (the residues of a matrix modulo 3, 2 taken as -1), rows in kernel order
// order[i].  Row sums s are kept as in build_block_tables: the byte
// 2 s + beta_i with beta_i = 128 or 126 by the parity of the row's number
// of nonzero entries d_i (s = d_i modulo 2), flip deltas 4 a_ij as signed
// bytes of one 32-bit word (a word sum or difference is exact because every
// resulting byte lies in [36, 220] for d_i <= 46), and the inner tables of
// the same form.  A row sum that is a multiple of 3 zeroes the term modulo
// 3, so every row is tested: the masks are 64 words per row (no overlap),
// entry e of row i (i < rows, i < n) having bit y set when the row sum at
// inner state y, O - 2 (sum of the row's entries in the inner columns y
// selects) for the outer sum O = (4 e - beta_i) / 2, is a multiple of 3.
inline void build_block_tables_mod3(const int64_t* M, int n, const std::vector<int>& order, int rows, int32_t* out) {
    const int W = byte_stride(n);
    std::vector<int64_t> words(block_table_words_mod3(n, rows), 0);
    auto put = [&](size_t word, int row, int64_t byte_value) { words[word] += byte_value << (8 * (row % 4)); };
    auto entry = [&](int i, int j) { return M[(size_t)order[i] * n + j]; };
    const size_t inner_at = (size_t)n * W, masks_at = inner_at + 12 * (size_t)W;
    std::vector<int64_t> correction(W, 0), low((size_t)W * 4, 0), high((size_t)W * 8, 0);
    for (int i = 0; i < 4 * W; ++i) {
        if (i >= n) {
            put((size_t)i / 4, i, 132);
            continue;
        }
        int64_t sum = 0, degree = 0;
        for (int j = 0; j < n; ++j) {
            sum += entry(i, j);
            degree += entry(i, j) != 0 ? 1 : 0;
        }
        const int64_t beta = degree % 2 == 0 ? 128 : 126;
        put((size_t)i / 4, i, 2 * sum + beta);
        for (int j = 0; j < n - 1; ++j) put((size_t)(1 + j) * W + i / 4, i, 4 * entry(i, j));
        correction[i / 4] += (128 - beta) << (8 * (i % 4));
        for (int a = 0; a < 4; ++a)
            low[(size_t)(i / 4) * 4 + a] += (4 * ((a & 1 ? entry(i, 0) : 0) + (a & 2 ? entry(i, 1) : 0))) << (8 * (i % 4));
        for (int b = 0; b < 8; ++b)
            high[(size_t)(i / 4) * 8 + b] +=
                (4 * ((b & 1 ? entry(i, 2) : 0) + (b & 2 ? entry(i, 3) : 0) + (b & 4 ? entry(i, 4) : 0)))
                << (8 * (i % 4));
        if (i < rows) {
            for (int e = 0; e < 64; ++e) {
                const int64_t outer = (4 * e - beta) / 2;
                uint32_t mask = 0;
                for (int y = 0; y < 32; ++y) {
                    int64_t inner = 0;
                    for (int j = 0; j < kBlockInnerBits; ++j)
                        if ((y >> j) & 1) inner += entry(i, j);
                    if ((outer - 2 * inner) % 3 == 0) mask |= 1u << y;
                }
                words[masks_at + (size_t)i * 64 + e] = mask;
            }
        }
    }
    for (int k = 0; k < W; ++k) {
        for (int a = 0; a < 4; ++a) words[inner_at + (size_t)k * 12 + a] = -(low[(size_t)k * 4 + a] / 2);
        for (int b = 0; b < 8; ++b)
            words[inner_at + (size_t)k * 12 + 4 + b] = -((high[(size_t)k * 8 + b] - correction[k]) / 2 - 0x40404040);
    }
    for (size_t k = 0; k < words.size(); ++k) out[k] = (int32_t)(uint32_t)(uint64_t)words[k];
}

}  // namespace phip
