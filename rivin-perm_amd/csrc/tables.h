// Host construction of the per-(matrix, modulus) tables read by glynn_chunk.
#pragma once
#include <stdint.h>

#include <vector>

#include "launchers.h"
#include "phip_arith.h"

namespace phip {

// Maximum absolute row sum of the centred residues of M modulo q.
inline uint64_t centred_row_bound(const int64_t* M, int n, int64_t q) {
    uint64_t S = 0;
    for (int r = 0; r < n; ++r) {
        uint64_t row = 0;
        for (int c = 0; c < n; ++c) {
            int64_t v = host_floor_mod(M[(size_t)r * n + c], q);
            if (v > q / 2) v -= q;
            row += (uint64_t)(v < 0 ? -v : v);
        }
        if (row > S) S = row;
    }
    return S;
}

// Largest group size whose exact products fit the kernel's bounds given the
// maximum absolute row sum S (S^G < 2^31, S^(G-1) < 2^23); 0 = wide rows.
inline int choose_group(uint64_t S) {
    const int b = S ? 64 - __builtin_clzll(S) : 0;
    if (b <= 7) return 4;
    if (b <= 10) return 3;
    if (b <= 15) return 2;
    return 0;
}

// Writes n * glynn_stride(n, group) int32 words: the row sums of state 0
// (every column sign +1), then the flip deltas 2a_ij of the n-1 Gray columns.
// group 0: residues in [0, q) one per word.  group > 0: exact centred values
// as 16-bit two's-complement halves, row 2k in the low half of word k.
inline void build_glynn_tables(const int64_t* M, int n, uint32_t q, int group, int32_t* out) {
    const int W = glynn_stride(n, group);
    std::vector<int64_t> v((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const int64_t r = host_floor_mod(M[(size_t)i * n + j], (int64_t)q);
            v[(size_t)i * n + j] = (group == 0 || r <= (int64_t)(q / 2)) ? r : r - (int64_t)q;
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
}

}  // namespace phip
