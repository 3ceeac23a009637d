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

// Greedy row order with a small live-column frontier for the frontier DP
// (perm_binary.py:282-386): a bounded number of starting rows (work budget
// 2,000,000 / (n^2 * ceil(average degree)), spread evenly over the rows
// ranked by (degree, index)); each step takes the row minimising (maximum
// width so far, new width, -closed columns, opened columns, row index); the
// best start minimises (maximum width, frontier area, order).  The matrix is
// given in CSR form (row r's columns are col_idx[row_ptr[r] .. row_ptr[r+1]),
// in any order, no repeats), so any order n works.  Returns 0 and writes the
// order, or -1.
extern "C" int phip_frontier_order(const int32_t* row_ptr, const int32_t* col_idx, int n, int32_t* order_out,
                                   int* max_width, int64_t* area) {
    if (n < 0) {
        phip::set_error("frontier order needs n >= 0");
        return -1;
    }
    if (!max_width || !area || (n > 0 && (!row_ptr || !col_idx || !order_out))) {
        phip::set_error("null argument");
        return -1;
    }
    *max_width = 0;
    *area = 0;
    if (n == 0) return 0;
    if (row_ptr[0] != 0) {
        phip::set_error("row_ptr must start at 0");
        return -1;
    }
    std::vector<std::vector<int>> cols(n);
    std::vector<int> degree(n, 0);
    std::vector<int> last_row(n, -1);
    int64_t entries = 0;
    for (int r = 0; r < n; ++r) {
        if (row_ptr[r + 1] < row_ptr[r]) {
            phip::set_error("row_ptr must be nondecreasing");
            return -1;
        }
        for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
            const int c = col_idx[e];
            if (c < 0 || c >= n) {
                phip::set_error("row mask contains a column outside the matrix");
                return -1;
            }
            if (last_row[c] == r) {
                phip::set_error("repeated column in a row");
                return -1;
            }
            last_row[c] = r;
            cols[r].push_back(c);
            ++degree[c];
            ++entries;
        }
    }
    const int64_t avg_ceil = std::max<int64_t>(1, (entries + n - 1) / n);
    const int64_t work_per_start = std::max<int64_t>(1, (int64_t)n * n * avg_ceil);
    const int start_count = (int)std::min<int64_t>(n, std::max<int64_t>(1, 2000000 / work_per_start));
    std::vector<int> ranked(n);
    for (int r = 0; r < n; ++r) ranked[r] = r;
    std::stable_sort(ranked.begin(), ranked.end(), [&](int a, int b) {
        return cols[a].size() != cols[b].size() ? cols[a].size() < cols[b].size() : a < b;
    });
    std::vector<int> starts;
    if (start_count == n) {
        starts = ranked;
    } else if (start_count == 1) {
        starts.push_back(ranked[0]);
    } else {
        for (int i = 0; i < start_count; ++i) starts.push_back(ranked[(int64_t)i * (n - 1) / (start_count - 1)]);
    }
    bool have_best = false;
    int best_width = 0;
    int64_t best_area = 0;
    std::vector<int> best_order;
    for (int start : starts) {
        std::vector<char> selected(n, 0);
        std::vector<int> processed(n, 0), order;
        int width = 0, maximum = 0;
        int64_t frontier_area = 0;
        for (int step = 0; step < n; ++step) {
            bool found = false;
            int b_max = 0, b_next = 0, b_negclosed = 0, b_opened = 0, b_row = 0;
            for (int row = (step == 0 ? start : 0); row < (step == 0 ? start + 1 : n); ++row) {
                if (selected[row]) continue;
                int opened = 0, closed = 0;
                for (int c : cols[row]) {
                    if (processed[c] == 0 && degree[c] > 1) ++opened;
                    if (processed[c] > 0 && processed[c] == degree[c] - 1) ++closed;
                }
                const int next = width + opened - closed;
                const int s0 = std::max(maximum, next);
                // Lexicographic (s0, next, -closed, opened, row).
                const bool better = !found || s0 < b_max ||
                                    (s0 == b_max && (next < b_next ||
                                    (next == b_next && (-closed < b_negclosed ||
                                    (-closed == b_negclosed && (opened < b_opened ||
                                    (opened == b_opened && row < b_row)))))));
                if (better) {
                    found = true;
                    b_max = s0;
                    b_next = next;
                    b_negclosed = -closed;
                    b_opened = opened;
                    b_row = row;
                }
            }
            if (!found) {
                phip::set_error("frontier ordering failed to select a row");
                return -1;
            }
            selected[b_row] = 1;
            order.push_back(b_row);
            for (int c : cols[b_row]) ++processed[c];
            width = b_next;
            maximum = std::max(maximum, width);
            frontier_area += width;
        }
        const bool better = !have_best || maximum < best_width ||
                            (maximum == best_width && (frontier_area < best_area ||
                            (frontier_area == best_area && order < best_order)));
        if (better) {
            have_best = true;
            best_width = maximum;
            best_area = frontier_area;
            best_order = order;
        }
    }
    for (int i = 0; i < n; ++i) order_out[i] = best_order[i];
    *max_width = best_width;
    *area = best_area;
    return 0;
}
