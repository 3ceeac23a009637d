// Host preprocessing of 0/1 matrices for the binary permanent path.  Host
// code only (no device call), so request admission never waits for Python
// loops on the serving path.
//
// The permanent of a 0/1 matrix with a perfect matching factors over the
// strongly connected components of its matching digraph (rows r -> r' when
// row r has a 1 in the column matched to r'); matrices without a perfect
// matching have permanent 0 (the decomposition the reference applies in
// perm_binary.py:609-752).  Here:
//   matching   Hopcroft-Karp: BFS layers from the free rows, then
//              vertex-disjoint shortest augmenting paths by an iterative
//              DFS that keeps each row's edge cursor for the whole phase;
//   components Tarjan's algorithm, iterative, one pass;
//   blocks     numbered by their smallest row; a block's rows and columns
//              keep ascending global order, which fixes the local column
//              labels of its row masks.
// The permanent does not depend on the block order or on the labels, only
// on the partition, which is unique.
//
// The frontier DP's row order (phip_frontier_order) starts from level
// structures of the row graph (rows adjacent when they share a column), as
// in bandwidth reduction: a pseudo-peripheral start row per connected
// component and Cuthill-McKee levels from it.  That order, its reverse and
// the identity are each refined by sifting (every row slides through all
// positions and stays at the best one), and the order with the least
// (maximum width, frontier area) is kept.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "device_context.h"
#include "permhip.h"

namespace {

constexpr int kInf = INT_MAX;

struct Csr {
    int n = 0;
    std::vector<int32_t> ptr, idx;
    int begin(int r) const { return ptr[r]; }
    int end(int r) const { return ptr[r + 1]; }
};

// Validates a CSR 0/1 matrix: ascending-free column lists in [0, n), no
// repeated column within a row.
int load_csr(const int32_t* row_ptr, const int32_t* col_idx, int n, Csr& g) {
    g.n = n;
    g.ptr.assign(row_ptr, row_ptr + n + 1);
    if (g.ptr[0] != 0) {
        phip::set_error("row_ptr must start at 0");
        return -1;
    }
    std::vector<int> last(n, -1);
    for (int r = 0; r < n; ++r) {
        if (g.ptr[r + 1] < g.ptr[r]) {
            phip::set_error("row_ptr must be nondecreasing");
            return -1;
        }
        for (int e = g.ptr[r]; e < g.ptr[r + 1]; ++e) {
            const int c = col_idx[e];
            if (c < 0 || c >= n) {
                phip::set_error("row mask contains a column outside the matrix");
                return -1;
            }
            if (last[c] == r) {
                phip::set_error("repeated column in a row");
                return -1;
            }
            last[c] = r;
        }
    }
    g.idx.assign(col_idx, col_idx + g.ptr[n]);
    return 0;
}

// Maximum matching; returns its size.
int hopcroft_karp(const Csr& g, std::vector<int>& match_row, std::vector<int>& match_col) {
    const int n = g.n;
    match_row.assign(n, -1);
    match_col.assign(n, -1);
    int size = 0;
    // Cheap start: each row takes its first free column.
    for (int r = 0; r < n; ++r) {
        for (int e = g.begin(r); e < g.end(r); ++e) {
            const int c = g.idx[e];
            if (match_col[c] < 0) {
                match_row[r] = c;
                match_col[c] = r;
                ++size;
                break;
            }
        }
    }
    std::vector<int> dist(n), queue(n), cursor(n), stack;
    stack.reserve(n);
    for (;;) {
        // Layers of alternating paths from the free rows.
        int head = 0, tail = 0;
        for (int r = 0; r < n; ++r) {
            if (match_row[r] < 0) {
                dist[r] = 0;
                queue[tail++] = r;
            } else {
                dist[r] = kInf;
            }
        }
        bool reachable_free_col = false;
        while (head < tail) {
            const int r = queue[head++];
            for (int e = g.begin(r); e < g.end(r); ++e) {
                const int owner = match_col[g.idx[e]];
                if (owner < 0) {
                    reachable_free_col = true;
                } else if (dist[owner] == kInf) {
                    dist[owner] = dist[r] + 1;
                    queue[tail++] = owner;
                }
            }
        }
        if (!reachable_free_col) break;
        for (int r = 0; r < n; ++r) cursor[r] = g.begin(r);
        // Vertex-disjoint shortest augmenting paths.  A row whose edges are
        // exhausted leaves the layered graph for the rest of the phase.
        for (int root = 0; root < n; ++root) {
            if (match_row[root] >= 0) continue;
            stack.clear();
            stack.push_back(root);
            bool augmented = false;
            while (!stack.empty() && !augmented) {
                const int u = stack.back();
                if (cursor[u] == g.end(u)) {
                    dist[u] = kInf;
                    stack.pop_back();
                    if (!stack.empty()) ++cursor[stack.back()];
                    continue;
                }
                const int c = g.idx[cursor[u]];
                const int owner = match_col[c];
                if (owner < 0) {
                    // Flip the path: each stacked row takes the column its
                    // cursor points at.
                    for (int v : stack) {
                        const int cv = g.idx[cursor[v]];
                        match_row[v] = cv;
                        match_col[cv] = v;
                    }
                    ++size;
                    augmented = true;
                } else if (dist[owner] != kInf && dist[owner] == dist[u] + 1) {
                    stack.push_back(owner);
                } else {
                    ++cursor[u];
                }
            }
        }
    }
    return size;
}

// Strongly connected components of the matching digraph; comp[r] in
// [0, count), numbered by smallest row.  Returns count.
int matching_components(const Csr& g, const std::vector<int>& match_col, std::vector<int>& comp) {
    const int n = g.n;
    std::vector<int> index(n, -1), low(n, 0), scc(n, -1), pending, frames, pos(n);
    std::vector<char> on_stack(n, 0);
    int counter = 0, found = 0;
    for (int root = 0; root < n; ++root) {
        if (index[root] >= 0) continue;
        frames.push_back(root);
        index[root] = low[root] = counter++;
        pos[root] = g.begin(root);
        pending.push_back(root);
        on_stack[root] = 1;
        while (!frames.empty()) {
            const int v = frames.back();
            if (pos[v] < g.end(v)) {
                const int w = match_col[g.idx[pos[v]++]];
                if (index[w] < 0) {
                    index[w] = low[w] = counter++;
                    pos[w] = g.begin(w);
                    pending.push_back(w);
                    on_stack[w] = 1;
                    frames.push_back(w);
                } else if (on_stack[w]) {
                    low[v] = std::min(low[v], index[w]);
                }
                continue;
            }
            frames.pop_back();
            if (!frames.empty()) low[frames.back()] = std::min(low[frames.back()], low[v]);
            if (low[v] == index[v]) {
                int w;
                do {
                    w = pending.back();
                    pending.pop_back();
                    on_stack[w] = 0;
                    scc[w] = found;
                } while (w != v);
                ++found;
            }
        }
    }
    // Renumber by smallest row.
    std::vector<int> relabel(found, -1);
    int next = 0;
    comp.assign(n, -1);
    for (int r = 0; r < n; ++r) {
        if (relabel[scc[r]] < 0) relabel[scc[r]] = next++;
        comp[r] = relabel[scc[r]];
    }
    return found;
}

// 1 with match_row / row_block / col_block / count filled, 0 without a
// perfect matching.
int decompose(const Csr& g, std::vector<int>& match_row, std::vector<int>& row_block, std::vector<int>& col_block,
              int& count) {
    std::vector<int> match_col;
    if (hopcroft_karp(g, match_row, match_col) != g.n) return 0;
    count = matching_components(g, match_col, row_block);
    col_block.assign(g.n, -1);
    for (int c = 0; c < g.n; ++c) col_block[c] = row_block[match_col[c]];
    return 1;
}

// --- frontier row order ---------------------------------------------------

struct FrontierCost {
    int width;
    int64_t area;
};

// Maximum and summed live-column count over the prefixes of an order: a
// column is live once one of its rows is processed and until its last row.
FrontierCost frontier_cost(const Csr& g, const std::vector<int>& degree, const std::vector<int>& order) {
    std::vector<int> seen(g.n, 0);
    int width = 0, widest = 0;
    int64_t area = 0;
    for (int r : order) {
        for (int e = g.begin(r); e < g.end(r); ++e) {
            const int c = g.idx[e];
            const int s = ++seen[c];
            if (s == 1 && degree[c] > 1) ++width;
            if (s == degree[c] && s > 1) --width;
        }
        widest = std::max(widest, width);
        area += width;
    }
    return {widest, area};
}

struct RowGraph {
    // Rows of each column (CSR by column).
    std::vector<int32_t> cptr, cidx;
};

// BFS levels over rows from `start`, restricted to rows with mark == 0;
// returns the visited rows in Cuthill-McKee order (neighbours of a row
// enter by ascending degree, ties by index) and the level count.  `mark`
// is left as given; `level` receives each visited row's level.
int cuthill_mckee(const Csr& g, const RowGraph& rg, int start, const std::vector<char>& mark,
                  std::vector<int>& level, std::vector<int>& out, std::vector<int>& stamp, int& epoch) {
    ++epoch;
    out.clear();
    out.push_back(start);
    stamp[start] = epoch;
    level[start] = 0;
    std::vector<int> nbrs;
    int levels = 1;
    for (size_t head = 0; head < out.size(); ++head) {
        const int r = out[head];
        nbrs.clear();
        for (int e = g.begin(r); e < g.end(r); ++e) {
            const int c = g.idx[e];
            for (int k = rg.cptr[c]; k < rg.cptr[c + 1]; ++k) {
                const int s = rg.cidx[k];
                if (mark[s] || stamp[s] == epoch) continue;
                stamp[s] = epoch;
                nbrs.push_back(s);
            }
        }
        std::sort(nbrs.begin(), nbrs.end(), [&](int a, int b) {
            const int da = g.end(a) - g.begin(a), db = g.end(b) - g.begin(b);
            return da != db ? da < db : a < b;
        });
        for (int s : nbrs) {
            level[s] = level[r] + 1;
            levels = std::max(levels, level[s] + 1);
            out.push_back(s);
        }
    }
    return levels;
}

std::vector<int> level_order(const Csr& g, const std::vector<int>& degree) {
    const int n = g.n;
    RowGraph rg;
    rg.cptr.assign(n + 1, 0);
    for (int c : g.idx) ++rg.cptr[c + 1];
    for (int c = 0; c < n; ++c) rg.cptr[c + 1] += rg.cptr[c];
    rg.cidx.resize(g.idx.size());
    {
        std::vector<int32_t> fill(rg.cptr.begin(), rg.cptr.end() - 1);
        for (int r = 0; r < n; ++r)
            for (int e = g.begin(r); e < g.end(r); ++e) rg.cidx[fill[g.idx[e]]++] = r;
    }
    (void)degree;
    std::vector<char> mark(n, 0);
    std::vector<int> level(n, 0), stamp(n, 0), comp, order;
    order.reserve(n);
    int epoch = 0;
    auto row_degree = [&](int r) { return g.end(r) - g.begin(r); };
    for (int seed = 0; seed < n; ++seed) {
        if (mark[seed]) continue;
        // Lowest-degree row of this component as the first guess.
        int levels = cuthill_mckee(g, rg, seed, mark, level, comp, stamp, epoch);
        int start = seed;
        for (int r : comp)
            if (row_degree(r) < row_degree(start) || (row_degree(r) == row_degree(start) && r < start)) start = r;
        levels = cuthill_mckee(g, rg, start, mark, level, comp, stamp, epoch);
        // Pseudo-peripheral row: move to a lowest-degree row of the last
        // level while that lengthens the level structure.
        for (int round = 0; round < 8; ++round) {
            int cand = -1;
            for (int r : comp)
                if (level[r] == levels - 1 &&
                    (cand < 0 || row_degree(r) < row_degree(cand) || (row_degree(r) == row_degree(cand) && r < cand)))
                    cand = r;
            std::vector<int> trial;
            const int trial_levels = cuthill_mckee(g, rg, cand, mark, level, trial, stamp, epoch);
            if (trial_levels <= levels) {
                // Restore the levels of the kept structure.
                cuthill_mckee(g, rg, start, mark, level, comp, stamp, epoch);
                break;
            }
            start = cand;
            levels = trial_levels;
            comp.swap(trial);
        }
        for (int r : comp) {
            mark[r] = 1;
            order.push_back(r);
        }
    }
    return order;
}

// Sifting (as for variable orders of decision diagrams): each row in turn
// slides through every position of the order by adjacent swaps and stays
// where (maximum width, area) is least.  A swap of neighbours a | b changes
// only the width after the prefix that ends between them, by the columns of
// a and of b that are not shared, so one slide costs O(deg) per position.
class Sifter {
public:
    Sifter(const Csr& g, const std::vector<int>& degree, std::vector<int>& order)
        : g_(g), degree_(degree), order_(order), n_(g.n), width_(g.n, 0), hist_(g.n + 2, 0), pre_(g.n, 0),
          stamp_(g.n, 0) {
        std::vector<int> seen(n_, 0);
        int w = 0;
        for (int k = 0; k < n_; ++k) {
            const int r = order_[k];
            for (int e = g_.begin(r); e < g_.end(r); ++e) {
                const int c = g_.idx[e];
                const int s = ++seen[c];
                if (s == 1 && degree_[c] > 1) ++w;
                if (s == degree_[c] && s > 1) --w;
            }
            width_[k] = w;
            ++hist_[w];
            area_ += w;
        }
        top_ = 0;
        for (int v = 0; v <= n_; ++v)
            if (hist_[v]) top_ = v;
    }

    FrontierCost cost() const { return {top_, area_}; }

    // One pass over the rows; returns true if the cost fell.
    bool pass(int64_t& budget) {
        const FrontierCost before = cost();
        std::vector<int> rows(order_);
        for (int r : rows) {
            if (budget <= 0) break;
            sift(r, budget);
        }
        const FrontierCost after = cost();
        return after.width < before.width || (after.width == before.width && after.area < before.area);
    }

private:
    bool live(int c, int count) const { return count > 0 && count < degree_[c]; }

    void set_width(int k, int w) {
        --hist_[width_[k]];
        ++hist_[w];
        area_ += w - width_[k];
        width_[k] = w;
        if (w > top_) top_ = w;
        while (top_ > 0 && hist_[top_] == 0) --top_;
    }

    // Swap the rows at positions k and k + 1; pre_ holds, for every column,
    // its rows at positions < k, and is left so.
    void swap_at(int k) {
        const int a = order_[k], b = order_[k + 1];
        ++epoch_;
        for (int e = g_.begin(b); e < g_.end(b); ++e) stamp_[g_.idx[e]] = epoch_;
        int delta = 0;
        for (int e = g_.begin(a); e < g_.end(a); ++e) {
            const int c = g_.idx[e];
            if (stamp_[c] == epoch_) continue;  // shared: unchanged
            delta += (int)live(c, pre_[c]) - (int)live(c, pre_[c] + 1);
        }
        ++epoch_;
        for (int e = g_.begin(a); e < g_.end(a); ++e) stamp_[g_.idx[e]] = epoch_;
        for (int e = g_.begin(b); e < g_.end(b); ++e) {
            const int c = g_.idx[e];
            if (stamp_[c] == epoch_) continue;
            delta += (int)live(c, pre_[c] + 1) - (int)live(c, pre_[c]);
        }
        order_[k] = b;
        order_[k + 1] = a;
        if (delta) set_width(k, width_[k] + delta);
    }

    void add_row(int r, int sign) {
        for (int e = g_.begin(r); e < g_.end(r); ++e) pre_[g_.idx[e]] += sign;
    }

    void sift(int r, int64_t& budget) {
        int pos = (int)(std::find(order_.begin(), order_.end(), r) - order_.begin());
        // pre_ = counts over positions < pos.
        std::fill(pre_.begin(), pre_.end(), 0);
        for (int k = 0; k < pos; ++k) add_row(order_[k], +1);
        FrontierCost best = cost();
        int best_pos = pos;
        auto consider = [&](int p) {
            const FrontierCost c = cost();
            if (c.width < best.width || (c.width == best.width && c.area < best.area)) {
                best = c;
                best_pos = p;
            }
        };
        const int deg = g_.end(r) - g_.begin(r);
        // Left to the front.
        while (pos > 0) {
            add_row(order_[pos - 1], -1);
            swap_at(pos - 1);
            --pos;
            budget -= deg + 4;
            consider(pos);
        }
        // Right to the back.
        while (pos < n_ - 1) {
            const int next = order_[pos + 1];
            swap_at(pos);
            add_row(next, +1);
            ++pos;
            budget -= deg + 4;
            consider(pos);
        }
        // Back to the best position.
        while (pos > best_pos) {
            add_row(order_[pos - 1], -1);
            swap_at(pos - 1);
            --pos;
        }
    }

    const Csr& g_;
    const std::vector<int>& degree_;
    std::vector<int>& order_;
    int n_;
    std::vector<int> width_, hist_, pre_, stamp_;
    int epoch_ = 0;
    int top_ = 0;
    int64_t area_ = 0;
};

}  // namespace

extern "C" int phip_binary_blocks(const uint64_t* masks, int n, int* block_count, int* block_sizes,
                                  uint64_t* block_masks) {
    if (n < 0 || n > 63) {
        phip::set_error("binary blocks need 0 <= n <= 63");
        return -1;
    }
    if (!block_count || (n > 0 && (!masks || !block_sizes || !block_masks))) {
        phip::set_error("null argument");
        return -1;
    }
    *block_count = 0;
    Csr g;
    g.n = n;
    g.ptr.assign(n + 1, 0);
    for (int r = 0; r < n; ++r) {
        if (masks[r] >> n) {
            phip::set_error("row mask has bits beyond the order");
            return -1;
        }
        for (uint64_t m = masks[r]; m; m &= m - 1) g.idx.push_back(__builtin_ctzll(m));
        g.ptr[r + 1] = (int32_t)g.idx.size();
    }
    std::vector<int> match_row, row_block, col_block;
    int count = 0;
    if (!decompose(g, match_row, row_block, col_block, count)) return 0;
    // Local column label = rank of the column among its block's columns.
    std::vector<int> local(n), fill(count, 0), rows_in(count, 0);
    for (int c = 0; c < n; ++c) local[c] = fill[col_block[c]]++;
    for (int r = 0; r < n; ++r) ++rows_in[row_block[r]];
    std::vector<int> offset(count + 1, 0);
    for (int b = 0; b < count; ++b) {
        offset[b + 1] = offset[b] + rows_in[b];
        block_sizes[b] = rows_in[b];
    }
    std::vector<int> cursor(offset.begin(), offset.end() - 1);
    for (int r = 0; r < n; ++r) {
        const int b = row_block[r];
        uint64_t m = 0;
        for (int e = g.begin(r); e < g.end(r); ++e) {
            const int c = g.idx[e];
            if (col_block[c] == b) m |= 1ull << local[c];
        }
        block_masks[cursor[b]++] = m;
    }
    *block_count = count;
    return 1;
}

extern "C" int phip_binary_blocks_csr(const int32_t* row_ptr, const int32_t* col_idx, int n, int32_t* row_match,
                                      int32_t* row_block, int32_t* col_block, int* block_count) {
    if (n < 0) {
        phip::set_error("binary blocks need n >= 0");
        return -1;
    }
    if (!block_count || (n > 0 && (!row_ptr || !row_block || !col_block))) {
        phip::set_error("null argument");
        return -1;
    }
    *block_count = 0;
    if (n == 0) return 1;
    Csr g;
    if (!col_idx && row_ptr[n] != 0) {
        phip::set_error("null argument");
        return -1;
    }
    static const int32_t kNoColumns = 0;
    if (load_csr(row_ptr, col_idx ? col_idx : &kNoColumns, n, g) != 0) return -1;
    std::vector<int> mr, rb, cb;
    int count = 0;
    if (!decompose(g, mr, rb, cb, count)) return 0;
    for (int i = 0; i < n; ++i) {
        if (row_match) row_match[i] = mr[i];
        row_block[i] = rb[i];
        col_block[i] = cb[i];
    }
    *block_count = count;
    return 1;
}

extern "C" int phip_frontier_order(const int32_t* row_ptr, const int32_t* col_idx, int n, int32_t* order_out,
                                   int* max_width, int64_t* area) {
    if (n < 0) {
        phip::set_error("frontier order needs n >= 0");
        return -1;
    }
    if (!max_width || !area || (n > 0 && (!row_ptr || !order_out))) {
        phip::set_error("null argument");
        return -1;
    }
    *max_width = 0;
    *area = 0;
    if (n == 0) return 0;
    if (!col_idx && row_ptr[n] != 0) {
        phip::set_error("null argument");
        return -1;
    }
    static const int32_t kNoColumns = 0;
    Csr g;
    if (load_csr(row_ptr, col_idx ? col_idx : &kNoColumns, n, g) != 0) return -1;
    std::vector<int> degree(n, 0);
    int64_t pair_work = 0;
    for (int c : g.idx) ++degree[c];
    for (int c = 0; c < n; ++c) pair_work += (int64_t)degree[c] * degree[c];
    std::vector<int> best(n);
    for (int r = 0; r < n; ++r) best[r] = r;
    FrontierCost best_cost = frontier_cost(g, degree, best);
    // The level structure walks every pair of rows sharing a column; dense
    // matrices (far outside the frontier DP's reach anyway) keep the
    // identity order.
    std::vector<std::vector<int>> starts;
    if (pair_work <= 50000000) {
        std::vector<int> cm = level_order(g, degree);
        starts.emplace_back(cm.rbegin(), cm.rend());
        starts.push_back(std::move(cm));
    }
    starts.push_back(best);
    // Refine every start by sifting until a pass gains nothing, within a
    // shared work budget of about 3e7 column updates (several passes at
    // n = 1000 with 3 entries a row), and keep the cheapest order.
    int64_t budget = 30000000;
    bool have = false;
    for (std::vector<int>& order : starts) {
        int64_t share = budget / 3;
        Sifter sifter(g, degree, order);
        for (int round = 0; round < 16 && share > 0; ++round)
            if (!sifter.pass(share)) break;
        const FrontierCost cost = sifter.cost();
        if (!have || cost.width < best_cost.width || (cost.width == best_cost.width && cost.area < best_cost.area)) {
            have = true;
            best_cost = cost;
            best = order;
        }
    }
    for (int i = 0; i < n; ++i) order_out[i] = best[i];
    *max_width = best_cost.width;
    *area = best_cost.area;
    return 0;
}
