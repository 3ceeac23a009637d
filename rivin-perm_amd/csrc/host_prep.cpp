// Host preprocessing of 0/1 matrices for the binary permanent path: a
// perfect matching and the irreducible blocks of the matching digraph
// (perm_binary.py:538-752; restated in permhip/binary_prep.py).  Host code
// only: no device call, so request admission does not wait for Python loops
// on the serving path.
//
// The traversal orders are those of binary_prep.py, so blocks come out in
// the same order and local column order:
//   matching  rows in order; BFS over augmenting paths, columns by
//             ascending index (lowest set bit first);
//   blocks    Kosaraju on row -> row edges r -> row_of_col[c] for each
//             column c of row r in ascending order: iterative DFS in row
//             order for the finish order, then the reversed graph (sources
//             listed in ascending order) by a LIFO stack from the last
//             finished row; each component's rows ascending.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "device_context.h"
#include "permhip.h"

namespace {

constexpr int kMaxRows = 63;

// match_row[r] = column of row r, or false when there is no perfect matching.
bool perfect_matching(const uint64_t* masks, int n, int* match_row) {
    int match_col[kMaxRows];
    for (int i = 0; i < n; ++i) match_row[i] = match_col[i] = -1;
    for (int start = 0; start < n; ++start) {
        int parent[kMaxRows];
        bool seen_row[kMaxRows] = {}, seen_col[kMaxRows] = {};
        int queue[kMaxRows];
        int head = 0, tail = 0, free_col = -1;
        seen_row[start] = true;
        queue[tail++] = start;
        while (head < tail && free_col < 0) {
            const int row = queue[head++];
            uint64_t avail = masks[row];
            while (avail) {
                const int col = __builtin_ctzll(avail);
                avail &= avail - 1;
                if (seen_col[col]) continue;
                seen_col[col] = true;
                parent[col] = row;
                const int owner = match_col[col];
                if (owner < 0) {
                    free_col = col;
                    break;
                }
                if (!seen_row[owner]) {
                    seen_row[owner] = true;
                    queue[tail++] = owner;
                }
            }
        }
        if (free_col < 0) return false;
        for (int col = free_col; col >= 0;) {
            const int row = parent[col];
            const int prev = match_row[row];
            match_row[row] = col;
            match_col[col] = row;
            col = prev;
        }
    }
    return true;
}

}  // namespace

extern "C" int phip_binary_blocks(const uint64_t* masks, int n, int* block_count, int* block_sizes,
                                  uint64_t* block_masks) {
    if (n < 0 || n > kMaxRows) {
        phip::set_error("binary blocks need 0 <= n <= 63");
        return -1;
    }
    if (!block_count || (n > 0 && (!masks || !block_sizes || !block_masks))) {
        phip::set_error("null argument");
        return -1;
    }
    *block_count = 0;
    for (int r = 0; r < n; ++r) {
        if (masks[r] >> n) {
            phip::set_error("row mask has bits beyond the order");
            return -1;
        }
    }
    int matching[kMaxRows];
    if (!perfect_matching(masks, n, matching)) return 0;
    if (n == 0) return 1;
    int row_of_col[kMaxRows];
    for (int r = 0; r < n; ++r) row_of_col[matching[r]] = r;
    std::vector<std::vector<int>> adj(n), rev(n);
    for (int r = 0; r < n; ++r) {
        for (uint64_t m = masks[r]; m; m &= m - 1) adj[r].push_back(row_of_col[__builtin_ctzll(m)]);
    }
    for (int s = 0; s < n; ++s)
        for (int t : adj[s]) rev[t].push_back(s);
    // First pass: finish order.
    std::vector<int> finish;
    std::vector<char> seen(n, 0);
    std::vector<std::pair<int, size_t>> stack;
    for (int start = 0; start < n; ++start) {
        if (seen[start]) continue;
        seen[start] = 1;
        stack.push_back({start, 0});
        while (!stack.empty()) {
            auto& top = stack.back();
            const int v = top.first;
            if (top.second == adj[v].size()) {
                stack.pop_back();
                finish.push_back(v);
                continue;
            }
            const int t = adj[v][top.second++];
            if (!seen[t]) {
                seen[t] = 1;
                stack.push_back({t, 0});
            }
        }
    }
    // Second pass on the reversed graph.
    std::fill(seen.begin(), seen.end(), 0);
    int written = 0, blocks = 0;
    std::vector<int> comp, lifo;
    for (auto it = finish.rbegin(); it != finish.rend(); ++it) {
        const int start = *it;
        if (seen[start]) continue;
        seen[start] = 1;
        comp.clear();
        lifo.assign(1, start);
        while (!lifo.empty()) {
            const int v = lifo.back();
            lifo.pop_back();
            comp.push_back(v);
            for (int t : rev[v]) {
                if (!seen[t]) {
                    seen[t] = 1;
                    lifo.push_back(t);
                }
            }
        }
        std::sort(comp.begin(), comp.end());
        const int k = (int)comp.size();
        for (int a = 0; a < k; ++a) {
            uint64_t local = 0;
            for (int b = 0; b < k; ++b)
                if ((masks[comp[a]] >> matching[comp[b]]) & 1u) local |= 1ull << b;
            block_masks[written++] = local;
        }
        block_sizes[blocks++] = k;
    }
    *block_count = blocks;
    return 1;
}
