// Host planner of the dense trellis DP in C++ (the steps of tropidist_amd/trellis.py, which restates
// cdist/trellis.py:47-193): rows over column positions, minimal span form, target-row reduction, op stream.
// Rows are bitsets over positions with their nonzero word span tracked, so the XORs of the span-form passes
// touch only the words a row occupies (rows of local codes are short intervals).
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <queue>
#include <utility>
#include <vector>

#include "../../include/tdp.h"

namespace tdp {
void set_error(const char *fmt, ...);
}

namespace {

struct Row {
    std::vector<uint64_t> w;
    int lo = 0, hi = 0;          // nonzero words lie in [lo, hi)

    bool zero() const { return lo >= hi; }
    void trim() {
        while (lo < hi && w[lo] == 0) lo++;
        while (hi > lo && w[hi - 1] == 0) hi--;
        if (lo >= hi) lo = hi = 0;
    }
    int lowbit() const { return zero() ? -1 : lo * 64 + __builtin_ctzll(w[lo]); }
    int bitlen() const { return zero() ? 0 : (hi - 1) * 64 + 64 - __builtin_clzll(w[hi - 1]); }
    bool test(int pos) const { return (w[pos >> 6] >> (pos & 63)) & 1; }
    void xor_with(const Row &o) {
        if (o.zero()) return;
        int a = zero() ? o.lo : std::min(lo, o.lo), b = zero() ? o.hi : std::max(hi, o.hi);
        for (int k = o.lo; k < o.hi; k++) w[k] ^= o.w[k];
        lo = a;
        hi = b;
        trim();
    }
};

// pass 1 keyed by start (the row ending earlier keeps the key), pass 2 keyed by end (the row starting earlier
// absorbs the other); dependent rows vanish
std::vector<Row> minimal_span_form(std::vector<Row> rows, int npos) {
    std::vector<int> by_start(npos + 1, -1);
    std::vector<Row> store;
    store.reserve(rows.size());
    for (Row &r : rows) {
        while (!r.zero()) {
            int s = r.lowbit();
            if (by_start[s] < 0) {
                by_start[s] = (int)store.size();
                store.push_back(std::move(r));
                break;
            }
            Row &held = store[by_start[s]];
            if (r.bitlen() < held.bitlen()) std::swap(r, held);
            r.xor_with(held);
        }
    }
    std::vector<Row *> stack;
    for (Row &r : store) stack.push_back(&r);
    std::sort(stack.begin(), stack.end(), [](Row *a, Row *b) { return a->lowbit() > b->lowbit(); });
    std::vector<int> by_end(npos + 2, -1);
    std::vector<Row> out;
    out.reserve(store.size());
    for (Row *rp : stack) {
        Row r = std::move(*rp);                  // store is not read again
        while (true) {
            int e = r.bitlen();
            if (by_end[e] < 0) {
                by_end[e] = (int)out.size();
                out.push_back(std::move(r));
                break;
            }
            Row &held = out[by_end[e]];
            if (held.lowbit() > r.lowbit()) {
                r.xor_with(held);
            } else {
                // held takes r, and r continues as held ^ r
                std::swap(held, r);
                r.xor_with(held);
            }
        }
    }
    // the reference keeps the order in which ends were first filled; out already is that order
    return out;
}

Row reduce_target_row(Row lrow, const std::vector<Row> &hrows, int npos) {
    std::vector<int> by_start(npos + 1, -1), by_end(npos + 2, -1);
    for (size_t i = 0; i < hrows.size(); i++) {
        by_start[hrows[i].lowbit()] = (int)i;
        by_end[hrows[i].bitlen()] = (int)i;
    }
    auto clear_start = [&](Row &x) {
        while (!x.zero()) {
            int q = by_start[x.lowbit()];
            if (q < 0) break;
            x.xor_with(hrows[q]);
        }
    };
    clear_start(lrow);
    while (!lrow.zero()) {
        int q = by_end[lrow.bitlen()];
        if (q < 0 || hrows[q].lowbit() < lrow.lowbit()) break;
        lrow.xor_with(hrows[q]);
        clear_start(lrow);
    }
    return lrow;
}

// Rows of the columns over positions of the order (trellis.rows_from_columns, cdist/trellis.py:47-57), trimmed.
int rows_over_positions(const uint64_t *h_words, int ncols, int nrows, const int32_t *order, std::vector<Row> &rows) {
    int rw = (nrows + 63) / 64, pw = (ncols + 63) / 64;
    rows.assign(nrows, Row());
    for (Row &r : rows) r.w.assign(pw > 0 ? pw : 1, 0);
    for (int pos = 0; pos < ncols; pos++) {
        int j = order[pos];
        if (j < 0 || j >= ncols) {
            tdp::set_error("order entry %d out of range", j);
            return TDP_EINVAL;
        }
        const uint64_t *c = h_words + (int64_t)j * rw;
        for (int k = 0; k < rw; k++)
            for (uint64_t x = c[k]; x; x &= x - 1) {
                int r = k * 64 + __builtin_ctzll(x);
                if (r < nrows) rows[r].w[pos >> 6] |= 1ull << (pos & 63);
            }
    }
    for (Row &r : rows) {
        r.lo = 0;
        r.hi = pw;
        r.trim();
    }
    return TDP_OK;
}

Row target_row(const uint8_t *targets, const int32_t *order, int ncols) {
    Row lrow;
    int pw = (ncols + 63) / 64;
    lrow.w.assign(pw > 0 ? pw : 1, 0);
    for (int pos = 0; pos < ncols; pos++)
        if (targets[order[pos]]) lrow.w[pos >> 6] |= 1ull << (pos & 63);
    lrow.lo = 0;
    lrow.hi = pw;
    lrow.trim();
    return lrow;
}

}  // namespace

extern "C" {

int tdp_plan_build(const uint64_t *h_words, int ncols, int nrows, const int32_t *order, const uint8_t *targets,
                   int nq, const int32_t *weights, uint64_t *ops, int64_t ops_cap, int64_t *nops, int32_t *peak,
                   double *work, int32_t *columns, int32_t *feasible) {
    if (ncols < 0 || nrows < 0 || nq < 0 || ops_cap < 0) {
        tdp::set_error("plan: bad sizes");
        return TDP_EINVAL;
    }
    int rw = (nrows + 63) / 64, pw = (ncols + 63) / 64;
    // rows over positions
    std::vector<Row> rows(nrows);
    for (Row &r : rows) r.w.assign(pw > 0 ? pw : 1, 0);
    for (int pos = 0; pos < ncols; pos++) {
        int j = order[pos];
        if (j < 0 || j >= ncols) {
            tdp::set_error("plan: order entry %d out of range", j);
            return TDP_EINVAL;
        }
        const uint64_t *c = h_words + (int64_t)j * rw;
        for (int k = 0; k < rw; k++)
            for (uint64_t x = c[k]; x; x &= x - 1) {
                int r = k * 64 + __builtin_ctzll(x);
                if (r < nrows) rows[r].w[pos >> 6] |= 1ull << (pos & 63);
            }
    }
    for (Row &r : rows) {
        r.lo = 0;
        r.hi = pw;
        r.trim();
    }
    std::vector<Row> hrows = minimal_span_form(std::move(rows), ncols);
    for (int q = 0; q < nq; q++) {
        uint64_t *qops = ops + (int64_t)q * ops_cap * 3;
        int32_t *qcols = columns + (int64_t)q * ops_cap;
        Row lrow;
        lrow.w.assign(pw > 0 ? pw : 1, 0);
        for (int pos = 0; pos < ncols; pos++)
            if (targets[(int64_t)q * ncols + order[pos]]) lrow.w[pos >> 6] |= 1ull << (pos & 63);
        lrow.lo = 0;
        lrow.hi = pw;
        lrow.trim();
        lrow = reduce_target_row(lrow, hrows, ncols);
        if (lrow.zero()) {
            feasible[q] = 0;
            nops[q] = 0;
            peak[q] = 0;
            work[q] = 0;
            continue;
        }
        feasible[q] = 1;
        std::vector<const Row *> all;
        for (const Row &r : hrows) all.push_back(&r);
        all.push_back(&lrow);
        std::vector<int> target(all.size(), 0);
        target.back() = 1;
        // rows opening / closing at each column, as index lists in row order (counting sort, no per-column lists)
        std::vector<int> open_at(ncols + 1, 0), close_at(ncols + 1, 0);
        for (size_t i = 0; i < all.size(); i++) {
            open_at[all[i]->lowbit() + 1]++;
            close_at[all[i]->bitlen()]++;
        }
        for (int j = 0; j < ncols; j++) {
            open_at[j + 1] += open_at[j];
            close_at[j + 1] += close_at[j];
        }
        std::vector<int> opening(all.size()), closing(all.size());
        {
            std::vector<int> fo(open_at.begin(), open_at.end() - 1), fc(close_at.begin(), close_at.end() - 1);
            for (size_t i = 0; i < all.size(); i++) {
                opening[fo[all[i]->lowbit()]++] = (int)i;
                closing[fc[all[i]->bitlen() - 1]++] = (int)i;
            }
        }
        std::vector<int> active, ending;
        std::vector<const uint64_t *> active_w;   // the active rows' words, beside `active`
        int64_t n = 0, k = 0;
        int bits = 0, pk = 0;
        double wk = 0;
        bool overflow = false;
        auto emit = [&](uint64_t code, uint64_t a, uint64_t b) {
            if (n >= ops_cap) { overflow = true; return; }
            qops[3 * n] = code;
            qops[3 * n + 1] = a;
            qops[3 * n + 2] = b;
            n++;
        };
        for (int j = 0; j < ncols && !overflow; j++) {
            for (int t = open_at[j]; t < open_at[j + 1]; t++) {
                active.push_back(opening[t]);
                active_w.push_back(all[opening[t]]->w.data());
                emit(0, 0, 0);
                bits++;
            }
            pk = std::max(pk, (int)active.size());
            uint64_t mask = 0;
            const int jw = j >> 6, jb = j & 63;
            uint64_t beyond = 0;                    // a row past slot 63 has this column (an error below)
            for (size_t p = 0; p < active_w.size(); p++) {
                const uint64_t bit = (active_w[p][jw] >> jb) & 1ull;
                if (p < 64) mask |= bit << p;
                else beyond |= bit;
            }
            if (mask || beyond) {
                if (active.size() > 63) {
                    tdp::set_error("plan: more than 63 active rows");
                    return TDP_EINVAL;
                }
                qcols[k++] = j;
                emit(1, mask, weights ? (uint64_t)weights[order[j]] : 1);
                wk += (double)(1ull << bits);
            }
            ending.clear();
            for (int t = close_at[j]; t < close_at[j + 1]; t++) {
                const int idx = closing[t];
                for (size_t p = 0; p < active.size(); p++)
                    if (active[p] == idx) ending.push_back((int)p);
            }
            std::sort(ending.rbegin(), ending.rend());
            for (int p : ending) {
                emit(2, (uint64_t)p, (uint64_t)target[active[p]]);
                active.erase(active.begin() + p);
                active_w.erase(active_w.begin() + p);
                bits--;
            }
        }
        if (overflow) {
            tdp::set_error("plan: op buffer of %lld entries too small", (long long)ops_cap);
            return TDP_ECAPACITY;
        }
        nops[q] = n;
        peak[q] = pk;
        work[q] = wk;
    }
    return TDP_OK;
}


// r[i] = GF(2) rank of the first i columns of the order (trellis.rank_profile, cdist/trellis.py:20-36).
// Columns are ncols bitsets of rw words (h_words, column-major); pivots are kept by one end bit with their word span.
// The rank does not depend on which end the pivots use, but the work does: a column reduced against pivots of
// earlier columns gains bits away from its pivot end, so the pivot end is chosen to point where the order is going
// (the lowest row when the columns' rows rise along the order, as in a banded matrix taken left to right, the highest
// when they fall), which keeps the cascade of reductions short.
}  // extern "C"

template <bool LOW>
static int rank_profile_by(const uint64_t *h_words, int ncols, int rw, const int32_t *order, int n, int32_t *out) {
    std::vector<Row> piv(64 * (size_t)rw);
    std::vector<char> has(64 * (size_t)rw, 0);
    int rank = 0;
    out[0] = 0;
    Row v;
    v.w.assign(rw, 0);
    for (int i = 0; i < n; i++) {
        const int j = order[i];
        if (j < 0 || j >= ncols) {
            tdp::set_error("tdp_rank_profile: order entry %d out of range", j);
            return TDP_EINVAL;
        }
        memcpy(v.w.data(), h_words + (size_t)j * rw, 8 * (size_t)rw);
        v.lo = 0;
        v.hi = rw;
        v.trim();
        while (!v.zero()) {
            const int key = LOW ? v.lowbit() : v.bitlen() - 1;
            if (!has[key]) {
                piv[key] = v;
                has[key] = 1;
                rank++;
                break;
            }
            v.xor_with(piv[key]);
        }
        out[i + 1] = rank;
    }
    return TDP_OK;
}

extern "C" {

int tdp_rank_profile(const uint64_t *h_words, int ncols, int rw, const int32_t *order, int n, int32_t *out) {
    if (!h_words || ncols < 0 || rw < 1 || !order || n < 0 || !out) {
        tdp::set_error("tdp_rank_profile: bad arguments");
        return TDP_EINVAL;
    }
    // direction of the order: mean lowest row of its first and last eighth of columns
    auto lowrow = [&](int j) -> double {
        if (j < 0 || j >= ncols) return 0.0;
        for (int k = 0; k < rw; k++) {
            uint64_t x = h_words[(size_t)j * rw + k];
            if (x) return k * 64 + __builtin_ctzll(x);
        }
        return 0.0;
    };
    const int m = std::max(1, n / 8);
    double head = 0, tail = 0;
    for (int i = 0; i < m && i < n; i++) {
        head += lowrow(order[i]);
        tail += lowrow(order[n - 1 - i]);
    }
    return head <= tail ? rank_profile_by<true>(h_words, ncols, rw, order, n, out)
                        : rank_profile_by<false>(h_words, ncols, rw, order, n, out);
}

}  // extern "C"

namespace {

// A query's rows with their spans and slots: the rows in minimal span form and the reduced target row (last), slots
// by an interval colouring of the closed spans (rows taken by start; a slot frees once its row has ended; the lowest
// free slot first).  Returns false when the target reduces to zero.
struct SlotRows {
    Row target;
    std::vector<const Row *> all;
    std::vector<int> st, en, slot;
    int ns = 0;
};

bool slot_rows(const std::vector<Row> &hrows, Row lrow_in, int ncols, SlotRows &t) {
    t.target = reduce_target_row(lrow_in, hrows, ncols);
    t.ns = 0;
    if (t.target.zero()) return false;
    t.all.clear();
    for (const Row &r : hrows) t.all.push_back(&r);
    t.all.push_back(&t.target);
    const int R = (int)t.all.size();
    t.st.assign(R, 0);
    t.en.assign(R, 0);
    t.slot.assign(R, 0);
    std::vector<int> idx(R);
    for (int r = 0; r < R; r++) {
        t.st[r] = t.all[r]->lowbit();
        t.en[r] = t.all[r]->bitlen() - 1;
        idx[r] = r;
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return t.st[a] < t.st[b]; });
    std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>, std::greater<std::pair<int, int>>>
        busy;
    std::priority_queue<int, std::vector<int>, std::greater<int>> freeq;
    for (int r : idx) {
        while (!busy.empty() && busy.top().first < t.st[r]) {
            freeq.push(busy.top().second);
            busy.pop();
        }
        int sl;
        if (!freeq.empty()) {
            sl = freeq.top();
            freeq.pop();
        } else {
            sl = t.ns++;
        }
        t.slot[r] = sl;
        busy.push({t.en[r], sl});
    }
    return true;
}

}  // namespace

extern "C" {

// The greedy straddle order of tropidist_amd/ordering.py: grow the processed set one column at a time, taking from
// the columns that share a row with a processed one the column that opens the fewest new rows net of the rows it
// closes, then the one touching the most open rows, then the lowest index; with no such column, the lowest
// unprocessed index.  Columns are ncols bitsets of ceil(nrows / 64) words.
int tdp_greedy_order(const uint64_t *h_words, int ncols, int nrows, int start, int32_t *order) {
    if (!h_words || !order || ncols < 0 || nrows < 0) {
        tdp::set_error("tdp_greedy_order: bad arguments");
        return TDP_EINVAL;
    }
    if (ncols == 0) return TDP_OK;
    const int rw = (nrows + 63) / 64 > 0 ? (nrows + 63) / 64 : 1;
    std::vector<std::vector<int>> col_rows(ncols), row_cols(nrows);
    for (int j = 0; j < ncols; j++)
        for (int k = 0; k < rw; k++)
            for (uint64_t x = h_words[(int64_t)j * rw + k]; x; x &= x - 1) {
                int r = k * 64 + __builtin_ctzll(x);
                if (r < nrows) {
                    col_rows[j].push_back(r);
                    row_cols[r].push_back(j);
                }
            }
    std::vector<int> remaining(nrows), in_front(ncols, 0), front;
    for (int r = 0; r < nrows; r++) remaining[r] = (int)row_cols[r].size();
    std::vector<char> opened(nrows, 0), done(ncols, 0);
    int current = (start >= 0 && start < ncols) ? start : 0, n = 0, next_free = 0;
    while (true) {
        order[n++] = current;
        done[current] = 1;
        for (int r : col_rows[current]) {
            opened[r] = 1;
            remaining[r]--;
            for (int j : row_cols[r])
                if (!done[j] && !in_front[j]) {
                    in_front[j] = 1;
                    front.push_back(j);
                }
        }
        if (in_front[current]) {
            in_front[current] = 0;
            front.erase(std::find(front.begin(), front.end(), current));
        }
        if (n == ncols) break;
        if (!front.empty()) {
            int best = -1, bnet = 0, btouch = 0;
            for (int j : front) {
                int nw = 0, closes = 0, touched = 0;
                for (int r : col_rows[j]) {
                    if (!opened[r]) nw += remaining[r] > 1;
                    else touched++;
                    if (remaining[r] == 1 && opened[r]) closes++;
                }
                int net = nw - closes;
                if (best < 0 || net < bnet || (net == bnet && (touched > btouch || (touched == btouch && j < best)))) {
                    best = j;
                    bnet = net;
                    btouch = touched;
                }
            }
            current = best;
        } else {
            while (done[next_free]) next_free++;
            current = next_free;
        }
    }
    return TDP_OK;
}

// The sparse trellis of each query (tropidist_amd/sparse_dp.Trellis, cdist/sparse_dp.py:63-145) in the word layout
// of the batched engine: minimal span form of the rows over positions, the reduced target row, an interval colouring
// of the closed row spans onto slots (rows by start; a slot frees once its row has ended; the lowest free slot first),
// then per position the column's slot mask and the closing masks and bits of both directions, and per cut the
// target row's slot while the row is active.  Query q's target column bits are targets[q * ncols + j].
// feasible[q] = 0 when the target reduces to zero; nslots[q] is the slot count, and the words are written only when
// it is at most max_slots (at most 120).  prm: nq x ncols x 10 words, tm: nq x (ncols + 1) x 2 words.
int tdp_sp_trellis(const uint64_t *h_words, int ncols, int nrows, const int32_t *order, const uint8_t *targets,
                   int nq, int max_slots, uint64_t *prm, uint64_t *tm, int32_t *nslots, int32_t *feasible) {
    if (!h_words || !order || !targets || !prm || !tm || !nslots || !feasible || ncols < 0 || nrows < 0 || nq < 0 ||
        max_slots < 1 || max_slots > 120) {
        tdp::set_error("tdp_sp_trellis: bad arguments");
        return TDP_EINVAL;
    }
    std::vector<Row> rows;
    int rc = rows_over_positions(h_words, ncols, nrows, order, rows);
    if (rc) return rc;
    std::vector<Row> hrows = minimal_span_form(std::move(rows), ncols);
    const int N = ncols;
    for (int q = 0; q < nq; q++) {
        SlotRows t;
        feasible[q] = slot_rows(hrows, target_row(targets + (int64_t)q * ncols, order, ncols), ncols, t) ? 1 : 0;
        nslots[q] = t.ns;
        if (!feasible[q] || t.ns > max_slots) continue;
        const int R = (int)t.all.size(), L = R - 1;
        uint64_t *P = prm + (int64_t)q * N * 10;
        uint64_t *T = tm + (int64_t)q * (N + 1) * 2;
        memset(P, 0, sizeof(uint64_t) * 10 * (size_t)N);
        memset(T, 0, sizeof(uint64_t) * 2 * (size_t)(N + 1));
        for (int r = 0; r < R; r++) {
            const int w = t.slot[r] >> 6;               // word of the slot bit: 0 = low, 1 = high
            const uint64_t b = 1ull << (t.slot[r] & 63);
            const Row &row = *t.all[r];
            for (int k = row.lo; k < row.hi; k++)
                for (uint64_t x = row.w[k]; x; x &= x - 1) P[(int64_t)(k * 64 + __builtin_ctzll(x)) * 10 + w] |= b;
            P[(int64_t)t.en[r] * 10 + 2 + w] |= b;       // forward closing mask
            P[(int64_t)t.st[r] * 10 + 6 + w] |= b;       // backward closing mask
            if (r == L) {
                P[(int64_t)t.en[r] * 10 + 4 + w] |= b;   // forward closing bits (the target row must be 1)
                P[(int64_t)t.st[r] * 10 + 8 + w] |= b;   // backward closing bits
                for (int c = t.st[r] + 1; c <= t.en[r]; c++) T[2 * c + w] |= b;
            }
        }
    }
    return TDP_OK;
}

// The lookahead lower bound of one query's sparse sweep, per cut (cdist/sparse_dp.py:148-192 defines it): the rows
// that still need an odd number of flips from the remaining columns -- the active rows, and the target row while it
// lies wholly on the remaining side -- each get the weight y_r = Q / (the most such rows any remaining column in the
// rows' window covers together with r), rounded down to a multiple of 1/Q, Q = 720720.  Every remaining column then
// carries total weight <= 1 (a dual solution of the covering LP), so ceil((extra + sum_r y_r [r odd]) / Q) faults
// are still needed.  Per cut c: ncls[c] classes (rows of equal weight merged into one slot mask, ascending weight) at
// masks + 2 * (c * cap + k) / weights + c * cap + k, and extra[c] = the target row's weight when it lies wholly
// ahead (else 0).  cap >= the slot count.
int tdp_sp_bounds(const uint64_t *h_words, int ncols, int nrows, const int32_t *order, const uint8_t *targets,
                  int forward, int cap, int32_t *ncls, int64_t *extra, uint64_t *masks, int64_t *weights) {
    if (!h_words || !order || !targets || !ncls || !extra || !masks || !weights || ncols < 0 || nrows < 0 ||
        cap < 1) {
        tdp::set_error("tdp_sp_bounds: bad arguments");
        return TDP_EINVAL;
    }
    std::vector<Row> rows;
    int rc = rows_over_positions(h_words, ncols, nrows, order, rows);
    if (rc) return rc;
    std::vector<Row> hrows = minimal_span_form(std::move(rows), ncols);
    SlotRows t;
    const int N = ncols;
    if (!slot_rows(hrows, target_row(targets, order, ncols), ncols, t)) {
        tdp::set_error("tdp_sp_bounds: the target reduces to zero (no query)");
        return TDP_EINVAL;
    }
    if (t.ns > cap) {
        tdp::set_error("tdp_sp_bounds: %d slots, room for %d classes", t.ns, cap);
        return TDP_EINVAL;
    }
    const int R = (int)t.all.size(), L = R - 1;
    const int64_t Q = 720720;
    std::vector<int> rel, cnt(N, 0);
    std::vector<std::pair<int64_t, int>> wr;     // (weight, slot)
    for (int c = 0; c <= N; c++) {
        rel.clear();
        bool lone = false;
        for (int r = 0; r < R; r++)
            if (t.st[r] < c && c <= t.en[r]) rel.push_back(r);
        if ((forward && t.st[L] >= c) || (!forward && t.en[L] < c)) {
            rel.push_back(L);
            lone = true;
        }
        ncls[c] = 0;
        extra[c] = 0;
        if (rel.empty()) continue;
        int lo = N, hi = 0;                      // the window [lo, hi) of remaining positions the rows reach
        if (forward) {
            lo = c;
            for (int r : rel) hi = std::max(hi, t.en[r] + 1);
        } else {
            hi = c;
            for (int r : rel) lo = std::min(lo, t.st[r]);
        }
        for (int j = lo; j < hi; j++) cnt[j] = 0;
        for (int r : rel)
            for (int j = std::max(lo, t.st[r]); j < std::min(hi, t.en[r] + 1); j++) cnt[j] += t.all[r]->test(j);
        wr.clear();
        for (int r : rel) {
            int most = 0;
            for (int j = std::max(lo, t.st[r]); j < std::min(hi, t.en[r] + 1); j++)
                if (t.all[r]->test(j)) most = std::max(most, cnt[j]);
            if (most < 1) {
                tdp::set_error("tdp_sp_bounds: row without a remaining column at cut %d", c);
                return TDP_EINVAL;
            }
            const int64_t w = Q / most;
            if (lone && r == L) extra[c] = w;
            else wr.push_back({w, t.slot[r]});
        }
        std::sort(wr.begin(), wr.end());
        int k = -1;
        for (size_t i = 0; i < wr.size(); i++) {
            if (i == 0 || wr[i].first != wr[i - 1].first) {
                k++;
                weights[(int64_t)c * cap + k] = wr[i].first;
                masks[2 * ((int64_t)c * cap + k)] = masks[2 * ((int64_t)c * cap + k) + 1] = 0;
            }
            masks[2 * ((int64_t)c * cap + k) + (wr[i].second >> 6)] |= 1ull << (wr[i].second & 63);
        }
        ncls[c] = k + 1;
    }
    return TDP_OK;
}

}  // extern "C"
