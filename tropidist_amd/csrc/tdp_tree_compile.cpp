// Compiler and runtime of the tree-realisation DP (host C++; the semantics of cdist/tree_dp.py:245-505).
//
// Given the columns of M (GF(2) vectors of rw words) and a binary tree over them, the DP keeps at every tree node S a
// table over the interface I_S = im(M_S) /\ im(M_rest): the least cost of a partial assignment of S's columns for
// each reachable boundary state.  The compiler computes, per node,
//   span_S      a basis of im(M_S), bottom up (the union of the children's spans);
//   outer_S     a basis of im(M_rest), top down (the parent's outer span and the sibling's span);
//   I_S         span_S /\ outer_S by the Zassenhaus sum-intersection elimination on doubled vectors [u | u], [w | 0];
// and per join P = A + B the generator maps of the compatible child pairs: the parent interface's basis vectors
// written as sums a + b of child interface vectors (the first parent_rank generators, one parent bit each) and the
// kernel of the child sum a + b = 0 (the free generators).  J = parent_rank + dim(I_A /\ I_B) must equal
// rank(M_A) + rank(M_B) + rank(M_rest) - rank(M); it is checked.  The execution order is a postorder whose child
// order, at each join, is the one giving the smaller peak of live table entries (both children's tables and the
// parent's coexist while joining).  Targets are solved for a particular solution e0 against an echelon basis of
// all columns that carries column-coefficient tags.
//
// The runtime walks the postorder on one stream: leaf tables, a fill and the joins (in Gray-counter chunks) of each
// internal node into the free arena space after the live tables, and a move of the parent down to its first child's
// offset (the stack compaction); the root's single entry is the answer.
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "../../include/tdp.h"

namespace tdp {
void set_error(const char *fmt, ...);
}

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

// An echelon set of GF(2) vectors keyed by their lowest set bit, each row `width` words, optionally carrying a tag of
// `tw` words (the combination of inputs it came from).
struct Echelon {
    int width = 0, tw = 0;
    std::vector<uint64_t> rows;       // nrows x (width + tw)
    std::vector<int> pivot;           // lowest set bit of each row
    std::vector<int> row_of_bit;      // bit -> row, -1

    void init(int w, int t) {
        width = w;
        tw = t;
        rows.clear();
        pivot.clear();
        row_of_bit.assign((size_t)w * 64, -1);
    }
    int stride() const { return width + tw; }
    size_t count() const { return pivot.size(); }
    const uint64_t *row(size_t i) const { return rows.data() + i * stride(); }

    static int lowest(const uint64_t *v, int width) {
        for (int k = 0; k < width; k++)
            if (v[k]) return k * 64 + __builtin_ctzll(v[k]);
        return -1;
    }
    // Reduces v (width + tw words) in place; returns the lowest bit left, -1 if v reduced to zero in its vector part.
    int reduce(uint64_t *v) const {
        for (;;) {
            int b = lowest(v, width);
            if (b < 0) return -1;
            int r = row_of_bit[b];
            if (r < 0) return b;
            const uint64_t *x = row((size_t)r);
            for (int k = b >> 6; k < stride(); k++) v[k] ^= x[k];
        }
    }
    // Adds a reduced, nonzero vector with lowest bit b.
    void add(const uint64_t *v, int b) {
        row_of_bit[b] = (int)pivot.size();
        pivot.push_back(b);
        rows.insert(rows.end(), v, v + stride());
    }
};

// A basis of the span of the given vectors (each `width` words), as a flat array.
void span_of(const std::vector<const uint64_t *> &vecs, int width, std::vector<uint64_t> &out) {
    Echelon e;
    e.init(width, 0);
    std::vector<uint64_t> v(width);
    for (const uint64_t *x : vecs) {
        std::copy(x, x + width, v.begin());
        int b = e.reduce(v.data());
        if (b >= 0) e.add(v.data(), b);
    }
    out = e.rows;
}

struct JoinRec {
    int lr = 0, rr = 0, pr = 0, J = 0;    // child ranks, parent rank, join dimension
    bool maps_ok = false;                 // every map fits one 64-bit word (ranks <= 63, J <= 62)
    std::vector<uint64_t> lm, rm;         // J generator maps (parent generators first)
};

struct TreeProgram {
    int ncols = 0, rw = 0, cw = 0, nnodes = 0;
    std::vector<uint64_t> cols;           // ncols x rw
    std::vector<int32_t> leaf, left, right;
    int matrix_rank = 0;
    std::vector<int32_t> rank;            // interface rank per node
    std::vector<std::vector<uint64_t>> iface;   // interface basis per node (rank x rw)
    std::vector<JoinRec> joins;           // per node (internal only)
    std::vector<int32_t> first, order;    // first child executed per internal node; execution order
    uint64_t peak_entries = 0;
    Echelon solver;                       // all columns, tags = column coefficients (cw words)
    uint64_t bytes = 0;                   // host bytes held by this object

    const uint64_t *col(int j) const { return cols.data() + (size_t)j * rw; }
};

uint64_t sat_add(uint64_t a, uint64_t b) { return a > UINT64_MAX - b ? UINT64_MAX : a + b; }
uint64_t pow2_sat(int r) { return r >= 64 ? UINT64_MAX : (1ull << r); }

// The interface span(U) /\ span(W) (both given as independent rows of rw words) by Zassenhaus: eliminate the doubled
// rows [u | u] and [w | 0] on their first half; a row whose first half vanishes leaves a vector of the intersection in
// its second half, and these vectors are independent (the kernel of the first half maps injectively onto them).
void intersect(const std::vector<uint64_t> &U, const std::vector<uint64_t> &W, int rw, std::vector<uint64_t> &out) {
    out.clear();
    const size_t nu = U.size() / (size_t)rw, nw = W.size() / (size_t)rw;
    if (!nu || !nw) return;
    Echelon e;
    e.init(rw, rw);                   // the "tag" half carries the second copy
    std::vector<uint64_t> v(2 * (size_t)rw);
    for (size_t i = 0; i < nw; i++) {
        std::copy(W.begin() + i * rw, W.begin() + (i + 1) * rw, v.begin());
        std::fill(v.begin() + rw, v.end(), 0);
        int b = e.reduce(v.data());
        if (b >= 0) e.add(v.data(), b);
    }
    for (size_t i = 0; i < nu; i++) {
        std::copy(U.begin() + i * rw, U.begin() + (i + 1) * rw, v.begin());
        std::copy(U.begin() + i * rw, U.begin() + (i + 1) * rw, v.begin() + rw);
        int b = e.reduce(v.data());
        if (b >= 0) e.add(v.data(), b);
        else out.insert(out.end(), v.begin() + rw, v.end());
    }
}

// Generator maps of a join: the parent basis vectors as a + b over the child interfaces, then the kernel of a + b.
// Returns false if a parent vector is not in the child sum (cannot happen for a consistent compilation).
bool join_maps(const std::vector<uint64_t> &A, const std::vector<uint64_t> &B, const std::vector<uint64_t> &P,
               int rw, JoinRec &jr) {
    const int x = (int)(A.size() / rw), y = (int)(B.size() / rw), k = (int)(P.size() / rw);
    const int tw = (x + y + 63) / 64 + 1;
    Echelon e;
    e.init(rw, tw);
    std::vector<uint64_t> v(rw + tw);
    std::vector<std::vector<uint64_t>> kernel;
    for (int i = 0; i < x + y; i++) {
        const uint64_t *src = i < x ? A.data() + (size_t)i * rw : B.data() + (size_t)(i - x) * rw;
        std::copy(src, src + rw, v.begin());
        std::fill(v.begin() + rw, v.end(), 0);
        v[rw + i / 64] |= 1ull << (i % 64);
        int b = e.reduce(v.data());
        if (b >= 0) e.add(v.data(), b);
        else kernel.emplace_back(v.begin() + rw, v.end());
    }
    jr.lr = x;
    jr.rr = y;
    jr.pr = k;
    jr.J = k + (int)kernel.size();
    jr.maps_ok = x <= 63 && y <= 63 && jr.J <= 62;
    if (!jr.maps_ok) return true;
    auto split = [&](const uint64_t *tag, uint64_t &l, uint64_t &r) {
        l = r = 0;
        for (int i = 0; i < x + y; i++)
            if ((tag[i / 64] >> (i % 64)) & 1) {
                if (i < x) l |= 1ull << i;
                else r |= 1ull << (i - x);
            }
    };
    jr.lm.assign(jr.J, 0);
    jr.rm.assign(jr.J, 0);
    for (int i = 0; i < k; i++) {
        std::copy(P.begin() + (size_t)i * rw, P.begin() + (size_t)(i + 1) * rw, v.begin());
        std::fill(v.begin() + rw, v.end(), 0);
        if (e.reduce(v.data()) >= 0) return false;
        split(v.data() + rw, jr.lm[i], jr.rm[i]);
    }
    for (size_t i = 0; i < kernel.size(); i++) split(kernel[i].data(), jr.lm[k + i], jr.rm[k + i]);
    return true;
}

int compile(TreeProgram &T, double max_seconds) {
    const auto t0 = Clock::now();
    auto late = [&]() { return max_seconds >= 0 && seconds_since(t0) > max_seconds; };
    const int n = T.nnodes, rw = T.rw;
    // the solver: echelon of all columns with their column tags
    T.solver.init(rw, T.cw);
    {
        std::vector<uint64_t> v(rw + T.cw);
        for (int j = 0; j < T.ncols; j++) {
            std::copy(T.col(j), T.col(j) + rw, v.begin());
            std::fill(v.begin() + rw, v.end(), 0);
            v[rw + j / 64] |= 1ull << (j % 64);
            int b = T.solver.reduce(v.data());
            if (b >= 0) T.solver.add(v.data(), b);
        }
        T.matrix_rank = (int)T.solver.count();
    }
    // spans bottom up
    std::vector<std::vector<uint64_t>> span(n);
    for (int i = 0; i < n; i++) {
        if (late()) return TDP_EDEADLINE;
        std::vector<const uint64_t *> in;
        if (T.leaf[i] >= 0) {
            in.push_back(T.col(T.leaf[i]));
        } else {
            for (int c : {T.left[i], T.right[i]})
                for (size_t r = 0; r < span[c].size() / rw; r++) in.push_back(span[c].data() + r * rw);
        }
        span_of(in, rw, span[i]);
    }
    std::vector<int> span_rank(n);
    for (int i = 0; i < n; i++) span_rank[i] = (int)(span[i].size() / rw);
    // outer spans top down (children after their parent: reverse postorder), interfaces
    std::vector<std::vector<uint64_t>> outer(n);
    std::vector<int> outer_rank(n, 0);
    T.iface.assign(n, {});
    T.rank.assign(n, 0);
    for (int i = n - 1; i >= 0; i--) {
        if (late()) return TDP_EDEADLINE;
        outer_rank[i] = (int)(outer[i].size() / rw);
        intersect(span[i], outer[i], rw, T.iface[i]);
        T.rank[i] = (int)(T.iface[i].size() / rw);
        if (T.leaf[i] < 0) {
            const int kids[2] = {T.left[i], T.right[i]};
            for (int s = 0; s < 2; s++) {
                const int c = kids[s], sib = kids[1 - s];
                std::vector<const uint64_t *> in;
                for (size_t r = 0; r < outer[i].size() / rw; r++) in.push_back(outer[i].data() + r * rw);
                for (size_t r = 0; r < span[sib].size() / rw; r++) in.push_back(span[sib].data() + r * rw);
                span_of(in, rw, outer[c]);
            }
        }
        std::vector<uint64_t>().swap(outer[i]);
    }
    // joins, the J identity, child order by peak live entries, execution order
    T.joins.assign(n, JoinRec());
    T.first.assign(n, -1);
    std::vector<uint64_t> peak(n, 0);
    for (int i = 0; i < n; i++) {
        if (late()) return TDP_EDEADLINE;
        const uint64_t size = pow2_sat(T.rank[i]);
        if (T.leaf[i] >= 0) {
            peak[i] = size;
            continue;
        }
        const int a = T.left[i], b = T.right[i];
        JoinRec &jr = T.joins[i];
        if (!join_maps(T.iface[a], T.iface[b], T.iface[i], rw, jr)) {
            tdp::set_error("tree compile: node %d's interface is not inside its children's sum", i);
            return TDP_EINVAL;
        }
        if (jr.J != span_rank[a] + span_rank[b] + outer_rank[i] - T.matrix_rank) {
            tdp::set_error("tree compile: node %d: join dimension %d breaks the rank identity", i, jr.J);
            return TDP_EINVAL;
        }
        const uint64_t sa = pow2_sat(T.rank[a]), sb = pow2_sat(T.rank[b]);
        const uint64_t all = sat_add(sat_add(sa, sb), size);
        const uint64_t ab = std::max(std::max(peak[a], sat_add(sa, peak[b])), all);
        const uint64_t ba = std::max(std::max(peak[b], sat_add(sb, peak[a])), all);
        T.first[i] = ab <= ba ? a : b;
        peak[i] = std::min(ab, ba);
    }
    T.peak_entries = n ? peak[n - 1] : 0;
    T.order.clear();
    if (n) {
        std::vector<std::pair<int, bool>> st{{n - 1, false}};
        while (!st.empty()) {
            auto [i, expanded] = st.back();
            st.pop_back();
            if (expanded || T.leaf[i] >= 0) {
                T.order.push_back(i);
                continue;
            }
            const int f = T.first[i], s = f == T.left[i] ? T.right[i] : T.left[i];
            st.push_back({i, true});
            st.push_back({s, false});
            st.push_back({f, false});
        }
    }
    uint64_t b = T.cols.size() * 8 + T.solver.rows.size() * 8 + (uint64_t)n * (sizeof(JoinRec) + 64);
    for (int i = 0; i < n; i++) b += T.iface[i].size() * 8 + T.joins[i].lm.size() * 16;
    T.bytes = b;
    return TDP_OK;
}

}  // namespace

extern "C" {

void *tdp_tree_compile(const uint64_t *cols, int ncols, int rw, const int32_t *node_leaf, const int32_t *node_left,
                       const int32_t *node_right, int nnodes, double max_seconds, int *status) {
    int dummy = 0;
    int &st = status ? *status : dummy;
    st = TDP_EINVAL;
    if (ncols < 0 || rw < 1 || nnodes < 0 || (ncols && !cols) || (nnodes && (!node_leaf || !node_left || !node_right))) {
        tdp::set_error("tree compile: bad arguments (ncols %d, rw %d, nodes %d)", ncols, rw, nnodes);
        return nullptr;
    }
    // the nodes must be a postorder of a binary tree using every column once
    std::vector<int> used(ncols, 0), parents(nnodes, 0);
    for (int i = 0; i < nnodes; i++) {
        if (node_leaf[i] >= 0) {
            if (node_leaf[i] >= ncols || used[node_leaf[i]]++) {
                tdp::set_error("tree compile: node %d: leaf %d repeated or out of range", i, node_leaf[i]);
                return nullptr;
            }
        } else if (node_left[i] < 0 || node_left[i] >= i || node_right[i] < 0 || node_right[i] >= i ||
                   node_left[i] == node_right[i] || parents[node_left[i]]++ || parents[node_right[i]]++) {
            tdp::set_error("tree compile: node %d: children must be distinct earlier nodes with one parent", i);
            return nullptr;
        }
    }
    if ((ncols && nnodes != 2 * ncols - 1) || (!ncols && nnodes) ||
        std::count(used.begin(), used.end(), 1) != ncols) {
        tdp::set_error("tree compile: %d nodes for %d columns", nnodes, ncols);
        return nullptr;
    }
    TreeProgram *T = new TreeProgram();
    T->ncols = ncols;
    T->rw = rw;
    T->cw = (ncols + 63) / 64 + 1;
    T->nnodes = nnodes;
    T->cols.assign(cols, cols + (size_t)ncols * rw);
    T->leaf.assign(node_leaf, node_leaf + nnodes);
    T->left.assign(node_left, node_left + nnodes);
    T->right.assign(node_right, node_right + nnodes);
    st = compile(*T, max_seconds);
    if (st != TDP_OK) {
        if (st == TDP_EDEADLINE) tdp::set_error("tree compilation exceeded max_compile_seconds");
        delete T;
        return nullptr;
    }
    return T;
}

void tdp_tree_free(void *h) { delete static_cast<TreeProgram *>(h); }

int tdp_tree_info(void *h, int64_t *out) {
    if (!h || !out) {
        tdp::set_error("tdp_tree_info: null argument");
        return TDP_EINVAL;
    }
    const TreeProgram &T = *static_cast<TreeProgram *>(h);
    int max_rank = 0, max_join = 0, all_maps = 1;
    for (int i = 0; i < T.nnodes; i++) {
        max_rank = std::max(max_rank, (int)T.rank[i]);
        if (T.leaf[i] < 0) {
            max_join = std::max(max_join, T.joins[i].J);
            all_maps &= T.joins[i].maps_ok ? 1 : 0;
        }
    }
    out[0] = T.matrix_rank;
    out[1] = (int64_t)std::min<uint64_t>(T.peak_entries, (uint64_t)INT64_MAX);
    out[2] = (int64_t)T.bytes;
    out[3] = max_rank;
    out[4] = max_join;
    out[5] = all_maps;
    return TDP_OK;
}

int tdp_tree_nodes(void *h, int32_t *rank, int32_t *order, int32_t *first, int32_t *joins) {
    if (!h || !rank || !order || !first || !joins) {
        tdp::set_error("tdp_tree_nodes: null argument");
        return TDP_EINVAL;
    }
    const TreeProgram &T = *static_cast<TreeProgram *>(h);
    for (int i = 0; i < T.nnodes; i++) {
        rank[i] = T.rank[i];
        order[i] = T.order[i];
        first[i] = T.first[i];
        const JoinRec &j = T.joins[i];
        const bool leaf = T.leaf[i] >= 0;
        joins[4 * i] = leaf ? -1 : j.lr;
        joins[4 * i + 1] = leaf ? -1 : j.rr;
        joins[4 * i + 2] = leaf ? -1 : j.pr;
        joins[4 * i + 3] = leaf ? -1 : j.J;
    }
    return TDP_OK;
}

int tdp_tree_maps(void *h, int node, uint64_t *lm, uint64_t *rm) {
    if (!h || !lm || !rm) {
        tdp::set_error("tdp_tree_maps: null argument");
        return TDP_EINVAL;
    }
    const TreeProgram &T = *static_cast<TreeProgram *>(h);
    if (node < 0 || node >= T.nnodes || T.leaf[node] >= 0 || !T.joins[node].maps_ok) {
        tdp::set_error("tdp_tree_maps: node %d has no 64-bit generator maps", node);
        return TDP_EINVAL;
    }
    const JoinRec &j = T.joins[node];
    std::copy(j.lm.begin(), j.lm.end(), lm);
    std::copy(j.rm.begin(), j.rm.end(), rm);
    return TDP_OK;
}

int tdp_tree_interface(void *h, int node, uint64_t *vecs) {
    if (!h || !vecs) {
        tdp::set_error("tdp_tree_interface: null argument");
        return TDP_EINVAL;
    }
    const TreeProgram &T = *static_cast<TreeProgram *>(h);
    if (node < 0 || node >= T.nnodes) {
        tdp::set_error("tdp_tree_interface: node %d out of range", node);
        return TDP_EINVAL;
    }
    std::copy(T.iface[node].begin(), T.iface[node].end(), vecs);
    return TDP_OK;
}

int tdp_tree_solve(void *h, const uint64_t *target, uint64_t *e0, int *feasible) {
    if (!h || !target || !e0 || !feasible) {
        tdp::set_error("tdp_tree_solve: null argument");
        return TDP_EINVAL;
    }
    const TreeProgram &T = *static_cast<TreeProgram *>(h);
    std::vector<uint64_t> v(T.rw + T.cw, 0);
    std::copy(target, target + T.rw, v.begin());
    *feasible = T.solver.reduce(v.data()) < 0;
    std::copy(v.begin() + T.rw, v.begin() + T.rw + (T.ncols + 63) / 64, e0);
    return TDP_OK;
}

int tdp_tree_run(void *h, const uint64_t *e0, int itemsize, void *d_arena, uint64_t arena_items,
                 uint64_t chunk_pairs, double max_seconds, uint64_t *value, uint64_t *stats, void *stream) {
    if (!h || !e0 || !d_arena || !value || !stats || !chunk_pairs ||
        (itemsize != 1 && itemsize != 2 && itemsize != 4 && itemsize != 8)) {
        tdp::set_error("tdp_tree_run: bad arguments");
        return TDP_EINVAL;
    }
    const TreeProgram &T = *static_cast<TreeProgram *>(h);
    const auto t0 = Clock::now();
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    auto late = [&]() { return max_seconds >= 0 && seconds_since(t0) > max_seconds; };
    auto deadline = [&]() {
        (void)hipStreamSynchronize(st);
        tdp::set_error("tree DP exceeded max_seconds; no exact result returned");
        return TDP_EDEADLINE;
    };
    if (T.nnodes == 0 || T.peak_entries > arena_items) {
        tdp::set_error("tdp_tree_run: arena of %llu items for a peak of %llu", (unsigned long long)arena_items,
                       (unsigned long long)T.peak_entries);
        return TDP_EINVAL;
    }
    const uint64_t inf = (uint64_t)T.ncols + 1;
    uint8_t *base = static_cast<uint8_t *>(d_arena);
    struct Live { int node; uint64_t off, size; };
    std::vector<Live> stack;
    uint64_t live = 0, seen_peak = 0, done = 0;
    for (int node : T.order) {
        if (late()) return deadline();
        const uint64_t size = 1ull << T.rank[node];
        seen_peak = std::max(seen_peak, live + size);
        if (live + size > arena_items) {
            tdp::set_error("tdp_tree_run: node %d's table would end past the arena", node);
            (void)hipStreamSynchronize(st);
            return TDP_EINVAL;
        }
        uint8_t *out = base + live * itemsize;
        if (T.leaf[node] >= 0) {
            const int j = T.leaf[node];
            const uint64_t bit = (e0[j / 64] >> (j % 64)) & 1;
            bool zero = true;
            for (int k = 0; k < T.rw; k++) zero = zero && T.col(j)[k] == 0;
            int rc = tdp_tree_leaf(itemsize, out, size, zero ? 0 : bit, bit ^ 1, inf, stream);
            if (rc) return rc;
            stack.push_back({node, live, size});
            live += size;
            continue;
        }
        Live second = stack.back();
        stack.pop_back();
        Live first = stack.back();
        stack.pop_back();
        const Live &L = first.node == T.left[node] ? first : second;
        const Live &R = first.node == T.left[node] ? second : first;
        const JoinRec &jr = T.joins[node];
        if (!jr.maps_ok) {
            tdp::set_error("tdp_tree_run: node %d needs join indices of more than 62 bits", node);
            return TDP_EINVAL;
        }
        int rc = tdp_tree_fill(itemsize, out, size, inf, stream);
        if (rc) return rc;
        const uint64_t work = 1ull << jr.J;
        for (uint64_t b = 0; b < work; b += std::min(chunk_pairs, work - b)) {
            if (late()) return deadline();
            const uint64_t e = b + std::min(chunk_pairs, work - b);
            rc = tdp_tree_join(itemsize, base + L.off * itemsize, base + R.off * itemsize, out, jr.lm.data(),
                               jr.rm.data(), (unsigned)jr.pr, (unsigned)jr.J, b, e, inf, stream);
            if (rc) return rc;
            done += e - b;
        }
        rc = tdp_tree_move(base + first.off * itemsize, out, size * itemsize, stream);
        if (rc) return rc;
        live = first.off + size;
        stack.push_back({node, first.off, size});
    }
    uint64_t root = 0;
    hipError_t e = hipMemcpyAsync(&root, base, (size_t)itemsize, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        tdp::set_error("tdp_tree_run: %s", hipGetErrorString(e));
        return TDP_EHIP;
    }
    if (late()) return deadline();
    if (live != 1 || stack.size() != 1 || seen_peak != T.peak_entries) {
        tdp::set_error("tdp_tree_run: arena schedule disagrees with the compiled peak");
        return TDP_EINVAL;
    }
    *value = root;
    stats[0] = done;
    stats[1] = seen_peak;
    return TDP_OK;
}

}  // extern "C"
