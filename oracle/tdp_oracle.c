/*
 * ORACLE -- test infrastructure only.  A plain, scalar CPU restatement of the reference's exact distance
 * kernels, used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.  Nothing in
 * the product (tropidist_amd/) links, loads or calls this file.
 *
 * Every function states the reference lines it follows.  The code is written for clarity (one scalar loop per
 * operation, no SIMD, no threads) so that it is easy to check against the reference's definitions:
 *
 *   dense trellis      cdist/viterbi.c:20-30 (saturating inc), 73-118 (dp_update), 122-138 (dp_expand_update),
 *                      155-172 (dp_select), 174-179 (dp_fill), 182-193 (dp_hist), 204-219 (dp_update_w)
 *   tree joins         cdist/tree_dp.c:13-39 (Gray-code join), 49-51 (tree_move)
 *   sparse tables      cdist/sparse_dp.c:21-44 (packed format, lookahead bound), 118-157 (sp_candidates),
 *                      162-202 (sp_merge), 207-236 (sp_join), 239-242 (sp_hist)
 *   MITM hash tables   cdist/gmitm.c:36-186 (Table, tab_create, tab_insert, tab_lookup, tab_build)
 *
 * Build: make -C oracle   (gcc -O2 -fPIC -shared tdp_oracle.c -o liboracle.so)
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_INF 255u

/* a + w saturating at 255: every entry is min(true value, 255) (viterbi.c:1-5, 20, 203-219) */
static inline uint8_t sat_add(uint8_t a, unsigned w) {
    unsigned s = (unsigned)a + w;
    return (uint8_t)(s > OR_INF ? OR_INF : s);
}

static inline uint8_t umin8(uint8_t a, uint8_t b) { return a < b ? a : b; }

static inline int top_bit(uint64_t x) { return 63 - __builtin_clzll(x); }

/* ------------------------------------------------------------------ dense trellis (viterbi.c) */

/* T[s] <- min(T[s], T[s^h] + w) for every s, from the old values (viterbi.c:73-118 unit, 204-219 weighted).
 * Each unordered pair {s, s^h} is visited once, from the member whose bit top_bit(h) is clear. */
void or_dp_update_w(uint8_t *T, uint64_t n, uint64_t h, int w) {
    if (h == 0 || n < 2) return;
    uint64_t hb = 1ULL << top_bit(h);
    for (uint64_t s = 0; s < n; s++) {
        if (s & hb) continue;
        uint64_t t = s ^ h;
        uint8_t a = T[s], b = T[t];
        T[s] = umin8(a, sat_add(b, (unsigned)w));
        T[t] = umin8(b, sat_add(a, (unsigned)w));
    }
}

void or_dp_update(uint8_t *T, uint64_t n, uint64_t h) { or_dp_update_w(T, n, h, 1); }

/* New top bit flipped by the current column: U[s ^ hl] = T[s] + w with U = T + n (viterbi.c:120-138). */
void or_dp_expand_update_w(uint8_t *T, uint64_t n, uint64_t hl, int w) {
    uint8_t *U = T + n;
    for (uint64_t s = 0; s < n; s++) U[s ^ hl] = sat_add(T[s], (unsigned)w);
}

void or_dp_expand_update(uint8_t *T, uint64_t n, uint64_t hl) { or_dp_expand_update_w(T, n, hl, 1); }

/* Keep the states whose bit `pos` equals `target` and drop that bit; returns n/2 (viterbi.c:149-172).
 * Output index k reads input index src(k) >= k, so an increasing in-place walk never reads a written entry. */
uint64_t or_dp_select(uint8_t *T, uint64_t n, int pos, int target) {
    uint64_t lowmask = (1ULL << pos) - 1;
    if ((n >> (pos + 1)) == 0) return n;          /* the reference's nb == 0 early return */
    uint64_t half = n >> 1;
    for (uint64_t k = 0; k < half; k++) {
        uint64_t src = ((k & ~lowmask) << 1) | ((uint64_t)target << pos) | (k & lowmask);
        T[k] = T[src];
    }
    return half;
}

void or_dp_fill(uint8_t *T, uint64_t n, uint8_t v) { memset(T, v, n); }      /* viterbi.c:174-179 */

void or_dp_hist(const uint8_t *T, uint64_t n, uint64_t *hist) {              /* viterbi.c:181-193 */
    for (int v = 0; v < 256; v++) hist[v] = 0;
    for (uint64_t s = 0; s < n; s++) hist[T[s]]++;
}

/* Whole plan executor over an encoded op stream (the loop of kernel.py:94-138, restated in C so the CPU
 * baseline pays no per-op Python cost).  Ops: code 0 = expand, 1 = update(mask, weight), 2 = select(pos, target),
 * each op is three uint64 words (code, a, b).  An expand immediately followed by an update whose top bit is the
 * new bit becomes one expand-update (kernel.py:115-121); otherwise the new half is filled with 255.
 * Returns the final entry T[0] (the caller applies final_value), or -1 on a malformed stream. */
int or_run_ops(const uint64_t *ops, int64_t nops, uint8_t *T, int peak_bits) {
    uint64_t n = 1, cap = 1ULL << peak_bits;
    T[0] = 0;
    for (int64_t i = 0; i < nops; i++) {
        uint64_t code = ops[3 * i], a = ops[3 * i + 1], b = ops[3 * i + 2];
        if (code == 0) {
            if (2 * n > cap) return -1;
            if (i + 1 < nops && ops[3 * (i + 1)] == 1 && (ops[3 * (i + 1) + 1] >> top_bit(n)) == 1) {
                or_dp_expand_update_w(T, n, ops[3 * (i + 1) + 1] ^ n, (int)ops[3 * (i + 1) + 2]);
                i++;
            } else {
                or_dp_fill(T + n, n, OR_INF);
            }
            n <<= 1;
        } else if (code == 1) {
            if (a >= n) return -1;
            or_dp_update_w(T, n, a, (int)b);
        } else if (code == 2) {
            if (n < 2) return -1;
            n = or_dp_select(T, n, (int)a, (int)b);
        } else {
            return -1;
        }
    }
    if (n != 1) return -1;
    return T[0];
}

/* ------------------------------------------------------------------ tree joins (tree_dp.c) */

/* For counters i in [start, stop): the Gray code g = i ^ (i >> 1) selects generators; the left/right indices are
 * the XOR of lm/rm over the set bits of g and the parent index is g restricted to the parent bits.
 * out[p] = min(out[p], L[l] + R[r]) when both are finite and the sum does not pass the infinity (tree_dp.c:13-39).
 * Restated as a direct recomputation per counter value; the result is independent of visiting order. */
#define OR_TREE_JOIN(T, NAME)                                                                          \
void NAME(const T *left, const T *right, T *out, const uint64_t *lm, const uint64_t *rm,                \
          unsigned parent_bits, unsigned generators, uint64_t start, uint64_t stop, uint64_t infinity) {  \
    (void)generators;                                                                                   \
    for (uint64_t i = start; i < stop; i++) {                                                           \
        uint64_t g = i ^ (i >> 1), li = 0, ri = 0, pi = 0;                                              \
        for (uint64_t rest = g; rest; rest &= rest - 1) {                                               \
            unsigned bit = (unsigned)__builtin_ctzll(rest);                                            \
            li ^= lm[bit];                                                                              \
            ri ^= rm[bit];                                                                              \
            if (bit < parent_bits) pi |= 1ULL << bit;                                                   \
        }                                                                                               \
        uint64_t a = left[li], b = right[ri];                                                           \
        if (a < infinity && b < infinity && a <= infinity - b && a + b < (uint64_t)out[pi])             \
            out[pi] = (T)(a + b);                                                                       \
    }                                                                                                   \
}
OR_TREE_JOIN(uint8_t, or_tree_join_u8)
OR_TREE_JOIN(uint16_t, or_tree_join_u16)
OR_TREE_JOIN(uint32_t, or_tree_join_u32)
OR_TREE_JOIN(uint64_t, or_tree_join_u64)

/* ------------------------------------------------------------------ sparse tables (sparse_dp.c) */

#define OR_KEYMASK ((1ULL << 56) - 1)
#define OR_VAL(x) ((int)((x) >> 56))
#define OR_QW 720720LL

typedef struct {
    int ncls;
    const uint64_t *mask;
    const int64_t *w;
    int64_t extra;
    uint64_t tmask;
    int K;
    int on;
} or_bound;

/* ceil((extra + sum_c w_c |(key ^ tmask) & mask_c|) / QW) (sparse_dp.c:35-40) */
static int or_hbound(const or_bound *b, uint64_t key) {
    uint64_t u = key ^ b->tmask;
    int64_t tot = b->extra;
    for (int c = 0; c < b->ncls; c++) tot += b->w[c] * (int64_t)__builtin_popcountll(u & b->mask[c]);
    return (int)((tot + OR_QW - 1) / OR_QW);
}

static int or_keep(const or_bound *b, uint64_t key, int v) { return !b->on || v + or_hbound(b, key) <= b->K; }

static int or_cmp_key(const void *pa, const void *pb) {
    uint64_t a = *(const uint64_t *)pa & OR_KEYMASK, b = *(const uint64_t *)pb & OR_KEYMASK;
    return a < b ? -1 : a > b;
}

/* Candidates s ^ upd (value + 1) of the states below the cap, filtered by the closing rows and the bound, sorted by
 * key (sparse_dp.c:116-157).  The candidate keys are distinct (XOR by upd is a bijection and the filter fixes the
 * cleared bits), so any sort by key gives the reference's array.  nbits, tmp and sort are accepted for signature
 * parity and unused. */
int64_t or_sp_candidates(const uint64_t *a, int64_t n, uint64_t upd, uint64_t sel_mask, uint64_t sel_bits, int cap,
                         int ncls, const uint64_t *mask, const int64_t *w, int64_t extra, uint64_t tmask, int K,
                         int use_h, int nbits, uint64_t *cand, uint64_t *tmp, int sort) {
    (void)nbits; (void)tmp; (void)sort;
    if (!upd) return 0;
    or_bound b = {ncls, mask, w, extra, tmask, K, use_h};
    uint64_t clr = ~sel_mask & OR_KEYMASK;
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) {
        int v = OR_VAL(a[i]);
        if (v >= cap) continue;
        uint64_t t = (a[i] & OR_KEYMASK) ^ upd;
        if ((t & sel_mask) != sel_bits) continue;
        t &= clr;
        if (!or_keep(&b, t, v + 1)) continue;
        cand[m++] = t | ((uint64_t)(v + 1) << 56);
    }
    qsort(cand, (size_t)m, sizeof(uint64_t), or_cmp_key);
    return m;
}

/* Merge of the filtered old table with the candidates (sparse_dp.c:159-202): keep the smaller value per key;
 * candidates that are new or lower go to ch; values >= kb leave the table, kb = min(keep_below or inf, cap + 1). */
int64_t or_sp_merge(const uint64_t *a, int64_t n, const uint64_t *cand, int64_t m, uint64_t sel_mask,
                    uint64_t sel_bits, int cap, int keep_below, int ncls, const uint64_t *mask, const int64_t *w,
                    int64_t extra, uint64_t tmask, int K, int use_h, uint64_t *out, uint64_t *ch, int64_t *nch) {
    or_bound b = {ncls, mask, w, extra, tmask, K, use_h};
    uint64_t clr = ~sel_mask & OR_KEYMASK;
    int kb = keep_below > 0 ? keep_below : (1 << 30);
    if (cap + 1 < kb) kb = cap + 1;
    int64_t i = 0, j = 0, no = 0, nc = 0;
    for (;;) {
        /* next old state that passes the closing-row filter */
        int have = 0;
        uint64_t ka = 0;
        int va = 0;
        while (i < n) {
            uint64_t k = a[i] & OR_KEYMASK;
            int v = OR_VAL(a[i]);
            i++;
            if ((k & sel_mask) == sel_bits) { ka = k & clr; va = v; have = 1; break; }
        }
        /* candidates with smaller keys than the old state come first */
        while (j < m && (!have || (cand[j] & OR_KEYMASK) < ka)) {
            uint64_t c = cand[j++];
            if (ch) ch[nc++] = c;
            if (OR_VAL(c) < kb) out[no++] = c;
        }
        if (!have) break;
        if (j < m && (cand[j] & OR_KEYMASK) == ka) {
            uint64_t c = cand[j++];
            int vc = OR_VAL(c);
            if (vc < va) {
                if (ch) ch[nc++] = c;
                if (vc < kb) out[no++] = c;
            } else if (va < kb && or_keep(&b, ka, va)) {
                out[no++] = ka | ((uint64_t)va << 56);
            }
        } else if (va < kb && or_keep(&b, ka, va)) {
            out[no++] = ka | ((uint64_t)va << 56);
        }
    }
    if (nch) *nch = nc;
    return no;
}

/* Least value(f) + value(b) over forward states f and backward states b with key(b) = key(f) ^ tmask
 * (sparse_dp.c:204-236), by binary search in the sorted backward table.  Ties keep the first f in the order the
 * reference visits them: the part with the tmask bit clear first, then the part with it set, each by key. */
int or_sp_join(const uint64_t *f, int64_t nf, const uint64_t *bt, int64_t nb, uint64_t tmask, uint64_t *fbest,
               uint64_t *bbest) {
    int best = -1;
    if (!nf || !nb) return -1;
    for (int part = 0; part < (tmask ? 2 : 1); part++) {
        for (int64_t i = 0; i < nf; i++) {
            uint64_t k = f[i] & OR_KEYMASK;
            if (tmask && (((k & tmask) != 0) != part)) continue;
            uint64_t q = k ^ tmask;
            int64_t lo = 0, hi = nb;                  /* first index with key >= q */
            while (lo < hi) {
                int64_t mid = (lo + hi) >> 1;
                if ((bt[mid] & OR_KEYMASK) < q) lo = mid + 1; else hi = mid;
            }
            if (lo < nb && (bt[lo] & OR_KEYMASK) == q) {
                int tot = OR_VAL(f[i]) + OR_VAL(bt[lo]);
                if (best < 0 || tot < best) { best = tot; *fbest = f[i]; *bbest = bt[lo]; }
            }
        }
    }
    return best;
}

void or_sp_hist(const uint64_t *a, int64_t n, int64_t *hist) {            /* sparse_dp.c:238-242 */
    for (int v = 0; v < 256; v++) hist[v] = 0;
    for (int64_t i = 0; i < n; i++) hist[OR_VAL(a[i])]++;
}

/* ------------------------------------------------------------------ MITM hash tables (gmitm.c:36-186) */

typedef struct {
    uint64_t mask;
    uint64_t *key;      /* 0 = empty; compact: bit63 | bits 16-62 of zl | 16-bit label; wide: bit63 | zl */
    uint64_t *lgs;      /* wide only: full label per slot */
    uint8_t *cost;      /* 255 = empty */
    int64_t count;
    int k;
} or_table;

#define OR_FP_MASK 0x7FFFFFFFFFFF0000ULL

void *or_tab_create(int log2slots, int wide) {
    or_table *t = (or_table *)calloc(1, sizeof(or_table));
    if (!t) return NULL;
    uint64_t n = 1ULL << log2slots;
    t->mask = n - 1;
    t->key = (uint64_t *)calloc(n, sizeof(uint64_t));
    t->cost = (uint8_t *)malloc(n);
    t->lgs = wide ? (uint64_t *)calloc(n, sizeof(uint64_t)) : NULL;
    if (!t->key || !t->cost || (wide && !t->lgs)) {
        free(t->key); free(t->cost); free(t->lgs); free(t);
        return NULL;
    }
    memset(t->cost, 255, n);
    return t;
}

void or_tab_free(void *p) {
    or_table *t = (or_table *)p;
    if (!t) return;
    free(t->key); free(t->cost); free(t->lgs); free(t);
}

int64_t or_tab_count(void *p) { return ((or_table *)p)->count; }

/* Linear probing from zh & mask; the entry with the same key (and, wide, the same label) keeps the least cost
 * (gmitm.c:94-134). */
static void or_tab_insert(or_table *t, uint64_t zh, uint64_t zl, uint64_t lg, uint8_t cost) {
    uint64_t i = zh & t->mask;
    uint64_t key = t->lgs ? (zl | (1ULL << 63)) : ((zl & OR_FP_MASK) | (lg & 0xFFFF) | (1ULL << 63));
    for (;;) {
        if (t->key[i] == 0) {
            t->key[i] = key;
            if (t->lgs) t->lgs[i] = lg;
            t->cost[i] = cost;
            t->count++;
            return;
        }
        if (t->key[i] == key && (!t->lgs || t->lgs[i] == lg)) {
            if (cost < t->cost[i]) t->cost[i] = cost;
            return;
        }
        i = (i + 1) & t->mask;
    }
}

/* Least cost <= R of an entry with the query's fingerprint and a different logical label, -1 if none
 * (gmitm.c:136-163). */
int or_tab_lookup(void *p, uint64_t zh, uint64_t zl, uint64_t lg, int R) {
    const or_table *t = (const or_table *)p;
    uint64_t i = zh & t->mask;
    int best = -1;
    for (;;) {
        uint64_t cur = t->key[i];
        if (cur == 0) return best;
        int match;
        if (t->lgs) match = cur == (zl | (1ULL << 63)) && t->lgs[i] != lg;
        else match = (cur & OR_FP_MASK) == (zl & OR_FP_MASK) && (cur & 0xFFFF) != (lg & 0xFFFF);
        if (match) {
            int c = t->cost[i];
            if (c <= R && (best < 0 || c < best)) best = c;
        }
        i = (i + 1) & t->mask;
    }
}

/* Insert the empty set and every subset of 1..k of the n columns with its size as the cost (gmitm.c:165-186),
 * enumerated as lexicographic index tuples. */
void or_tab_build(void *p, int n, const uint64_t *zh, const uint64_t *zl, const uint64_t *lg, int k) {
    or_table *t = (or_table *)p;
    t->k = k;
    or_tab_insert(t, 0, 0, 0, 0);
    if (k < 1 || n < 1) return;
    int idx[64];
    uint64_t ah[65], al[65], ag[65];
    ah[0] = al[0] = ag[0] = 0;
    int depth = 0;
    idx[0] = -1;
    /* iterative DFS over increasing index tuples of length 1..k */
    for (;;) {
        idx[depth]++;
        if (idx[depth] >= n) {
            if (depth == 0) break;
            depth--;
            continue;
        }
        int j = idx[depth];
        ah[depth + 1] = ah[depth] ^ zh[j];
        al[depth + 1] = al[depth] ^ zl[j];
        ag[depth + 1] = ag[depth] ^ lg[j];
        or_tab_insert(t, ah[depth + 1], al[depth + 1], ag[depth + 1], (uint8_t)(depth + 1));
        if (depth + 1 < k && depth + 1 < 64) {
            depth++;
            idx[depth] = j;
        }
    }
}
