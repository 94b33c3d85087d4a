// Weight-bounded exact search on gfx950: column sets e with H e = 0, L e != 0 and |e| <= K, by the canonical
// anchored search of cdist/gmitm.c:188-1023 (the engine behind cdist/search.py:145-441).
//
// The search tree is the reference's: a node is a prefix A (the anchor, then one column per level) with its residual
// syndrome; its children are the columns of the residual row with the fewest free columns, each child forbidding the
// siblings before it; when the remaining budget R = K - |A| reaches the depth k of the exact table, one lookup
// decides whether a completion can exist, and every hit is re-derived by an explicit completion search.  Optional
// prunings are the reference's: relaxed tables, the covering bound with its dual ascent, exclusion pairs, kernel
// dominance sets and the pattern budget.
//
// Device design.  The DFS runs level by level: one node buffer per depth (struct of arrays, word w of node i at
// base[w * cap + i]), and every launch expands a chunk of nodes of one depth, one thread per node, into the next
// depth's buffer.  The host always expands the deepest non-empty depth (refilling it from the depth above when it
// runs short), so the buffers stay bounded by depth x chunk x row degree.  A node carries its full forbidden set as
// a bitmask instead of the reference's per-thread stamps and undo lists: a child's set is its parent's plus the
// siblings before it, the exclusion partners and the kernel sets its column closes, which is the set the
// reference's stamps hold at that node.  Nodes whose children are decided by one lookup each (R - 1 <= k) probe
// their children in place and store only those that hit; hits are completed in the thread that finds them.
//
// The visited nodes, hits, false hits and pruning counts per node are the reference's (its per-depth node counts
// are a function of the tree alone); solutions come in another order (callers sort them).
#include "tdp_tables.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <mutex>
#include <vector>

namespace tdp {
namespace {

constexpr int kDmax = 40;          // deepest threshold (cdist/gmitm.c:31)
constexpr int kTmax = 4;           // lookup tables per search
constexpr int kCvMax = 512;        // defect rows the dual ascent handles
constexpr int kInvMax = 1024;
constexpr int kBlock = 128;
constexpr int kIncLdsBytes = 32 << 10;   // incidence arrays staged in LDS up to this size
constexpr int kStageWords = 24;           // syndrome + forbidden words staged in LDS up to this many
constexpr int kCompleteMax = 8;    // completion depth: the deepest table

// per-launch counters, summed into the run's statistics
enum { S_LOOKUPS, S_HITS, S_FALSE, S_TASKS, S_CPRUNED, S_APRUNED, S_CCALLS, S_NODES_A, S_NODES_A1, S_N };
// run statistics on the device: nodes per depth [0, kDmax + 2), then the counters of cdist/gmitm.c:857-869
constexpr int kStatBase = kDmax + 2;
// Statistics words sit 128 bytes apart: every block adds its counts with one atomic each, and atomics on one cache
// line serialise (about 12 ns each on MI355X).
constexpr int kStatStride = 16;

struct GmFlags {
    int stop;
    int overflow;
    int best;                      // INT_MAX: none
    unsigned nsol;                 // solutions claimed (stored while below the cap)
    unsigned long long nfound;
    unsigned count[kDmax + 2];     // nodes held per depth
};

struct Level {
    uint64_t *base;
    uint32_t cap;
};

struct GmDev {
    int W, N, M, wmax, K, td, enum_all, kernel_mode, lw, cover, ascent, ntab, main_tab;
    int nnz, inc_lds, inc_words, stage;
    int cvmax;                     // defect rows per slot of the ascent arrays: min(kCvMax, wmax * K) >= pc
    int off_hdr, off_path, off_s, off_zh, off_zl, off_lg, off_forb, off_used;
    double inv_wmax;
    const uint64_t *col, *lg, *full_lg;
    const int *row_start, *row_cols, *col_start, *col_rows;
    const uint64_t *tzh[kTmax], *tzl[kTmax];
    TabView tab[kTmax];
    int tab_wide[kTmax], tab_k[kTmax];
    const int *ex_start, *ex_cols;
    int nks;
    const int *ks_start, *ks_cols, *kc_start, *kc_ids;
    const int *cls;
    const uint8_t *onpat;
    int slack;
    const double *inv;             // inv[m] = 1 / m, 1 <= m <= kInvMax
    int sol_cap, sol_stride;
    int *sol;
    // dual-ascent scratch per thread slot
    double *cv_load;
    unsigned *cv_stamp, *cv_id;
    int *cv_drow, *cv_av, *cv_ord;
    double *cv_dy;
};

#define NODE(L, w, i) ((L).base[(size_t)(w) * (L).cap + (i)])

__device__ __forceinline__ bool stopped(const GmFlags *fl) {
    return __hip_atomic_load(&fl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// What one thread reads of the node it visits: its syndrome and forbidden words (staged in LDS, or in the level
// buffer), the row/column incidence (LDS or global), and the node's index for the rest of its words.
struct Ctx {
    const uint64_t *s;
    unsigned ss;                   // syndrome word w at s[w * ss]
    const uint64_t *fb;
    unsigned fs;                   // forbidden word w at fb[w * fs]
    const int *rs, *rc, *cs, *cr;  // row_start, row_cols, col_start, col_rows
    uint32_t i;
};

__device__ __forceinline__ uint64_t syn(const Ctx &C, int w) { return C.s[(size_t)w * C.ss]; }

__device__ __forceinline__ uint64_t forb_word(const Ctx &C, int w) { return C.fb[(size_t)w * C.fs]; }

__device__ __forceinline__ bool is_forb(const Ctx &C, int f) { return (forb_word(C, f >> 6) >> (f & 63)) & 1; }

__device__ __forceinline__ int path_at(const GmDev &P, const Level &L, uint32_t i, int p) {
    return (int)((NODE(L, P.off_path + (p >> 2), i) >> ((p & 3) * 16)) & 0xFFFF);
}

__device__ __forceinline__ int free_in_row(const Ctx &C, int r) {
    int av = 0;
    for (int q = C.rs[r]; q < C.rs[r + 1]; q++) av += !is_forb(C, C.rc[q]);
    return av;
}

// 0: over the pattern budget; 1: takes the free slot of its class; 2: counts as an extra (cdist/gmitm.c:347-365)
__device__ __forceinline__ int pat_how(const GmDev &P, const Level &L, uint32_t i, int f) {
    if (!P.cls) return 1;
    int cl = P.cls[f];
    if (cl >= 0 && P.onpat[cl] && !((NODE(L, P.off_used + (cl >> 6), i) >> (cl & 63)) & 1)) return 1;
    int extras = (int)(NODE(L, P.off_hdr, i) & 0xFFFF);
    return extras < P.slack ? 2 : 0;
}

// is the logical value of path(0..a) + B(0..nb) nonzero (g: the labelled-mode running label)
__device__ bool sol_nonzero(const GmDev &P, const Level &L, uint32_t i, int a, const int *B, int nb, uint64_t g) {
    if (!P.full_lg) return g != 0;
    for (int w = 0; w < P.lw; w++) {
        uint64_t x = 0;
        for (int p = 0; p < a; p++) x ^= P.full_lg[(size_t)path_at(P, L, i, p) * P.lw + w];
        for (int b = 0; b < nb; b++) x ^= P.full_lg[(size_t)B[b] * P.lw + w];
        if (x) return true;
    }
    return false;
}

__device__ void record(const GmDev &P, GmFlags *fl, const Level &L, uint32_t i, int a, const int *B, int nb) {
    const int weight = a + nb;
    atomicMin(&fl->best, weight);
    atomicAdd(&fl->nfound, 1ull);
    unsigned idx = atomicAdd(&fl->nsol, 1u);
    if ((int)idx < P.sol_cap) {
        int *dst = P.sol + (size_t)idx * P.sol_stride;
        dst[0] = weight;
        for (int q = 0; q < weight && q + 1 < P.sol_stride; q++) dst[q + 1] = q < a ? path_at(P, L, i, q) : B[q - a];
    }
    if (!P.enum_all) __hip_atomic_store(&fl->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every B (|B| <= R) that cancels the node's syndrome, with a nonzero total label, disjoint from the node's path
// (cdist/gmitm.c:444-482): branch on the lowest residual row; a column taken at a level stays excluded for its
// later siblings and their subtrees.  Returns the number of completions recorded.
__device__ int complete(const GmDev &P, const Level &L, const Ctx &C, int a, int R, GmFlags *fl) {
    const uint32_t i = C.i;
    int B[kCompleteMax + 1], rowB[kCompleteMax + 1], posB[kCompleteMax + 1];
    int found = 0, lev = 0;
    bool enter = true;
    for (;;) {
        if (enter) {
            enter = false;
            bool descend = false;
            int pc = 0, r = -1;
            for (int w = 0; w < P.W; w++) {
                uint64_t x = syn(C, w);
                for (int b = 0; b < lev; b++) x ^= P.col[(size_t)B[b] * P.W + w];
                pc += __popcll(x);
                if (r < 0 && x) r = w * 64 + __builtin_ctzll(x);
            }
            if (pc == 0) {
                uint64_t g = NODE(L, P.off_lg, i);
                for (int b = 0; b < lev; b++) g ^= P.lg[B[b]];
                if (sol_nonzero(P, L, i, a, B, lev, g) != (bool)P.kernel_mode) {
                    record(P, fl, L, i, a, B, lev);
                    found++;
                }
            } else if (R - lev > 0 && pc <= P.wmax * (R - lev)) {
                rowB[lev] = r;
                posB[lev] = C.rs[r] - 1;
                descend = true;
            }
            if (!descend) {
                if (lev == 0) return found;
                lev--;
            }
        }
        if (stopped(fl)) return found;
        const int r = rowB[lev], end = C.rs[r + 1];
        int q = posB[lev] + 1, f = -1;
        for (; q < end; q++) {
            const int c = C.rc[q];
            bool excluded = false;
            for (int p = 0; p < a && !excluded; p++) excluded = path_at(P, L, i, p) == c;
            for (int l = 0; l < lev && !excluded; l++)
                for (int t = C.rs[rowB[l]]; t <= posB[l]; t++)
                    if (C.rc[t] == c) {
                        excluded = true;
                        break;
                    }
            if (!excluded) {
                f = c;
                break;
            }
        }
        if (f < 0) {
            if (lev == 0) return found;
            lev--;
            continue;
        }
        posB[lev] = q;
        B[lev] = f;
        lev++;
        enter = true;
    }
}

// Covering bound (cdist/gmitm.c:484-596): a feasible dual y of the covering LP of the residual rows by the free
// columns; true if it exceeds R.  Same operations in the same order as the reference, in double precision.
__device__ bool cover_prune(const GmDev &P, const Ctx &C, int R, int pc, uint32_t slot, unsigned &apruned) {
    double sum = 0.0;
    const double lim = R + 1e-9, inv_wmax = P.inv_wmax;
    int left = pc;
    const bool ascent = P.ascent && pc <= kCvMax;
    int *drow = P.cv_drow + (size_t)slot * P.cvmax;
    double *dy = P.cv_dy + (size_t)slot * P.cvmax;
    int nd = 0;
    if (left * inv_wmax > lim) return true;
    for (int w = 0; w < P.W; w++) {
        uint64_t x = syn(C, w);
        while (x) {
            const int r = w * 64 + __builtin_ctzll(x);
            x &= x - 1;
            left--;
            int m = 0;
            for (int q = C.rs[r]; q < C.rs[r + 1]; q++) {
                const int f = C.rc[q];
                if (is_forb(C, f)) continue;
                int o = 0;
                for (int t = C.cs[f]; t < C.cs[f + 1]; t++) {
                    const int rr = C.cr[t];
                    o += (int)((syn(C, rr >> 6) >> (rr & 63)) & 1);
                }
                m = o > m ? o : m;
            }
            if (m == 0) return true;
            sum += P.inv[m];
            if (sum + left * inv_wmax > lim) return true;
            if (ascent) {
                drow[nd] = r;
                dy[nd] = P.inv[m];
                nd++;
            } else if (sum + left <= lim) {
                return false;
            }
        }
    }
    if (!ascent) return false;
    double *load = P.cv_load + (size_t)slot * P.N;
    unsigned *stamp = P.cv_stamp + (size_t)slot * P.N;
    unsigned id = ++P.cv_id[slot];
    if (id == 0) {                                   // wrapped: no stale stamp may equal a new id
        for (int f = 0; f < P.N; f++) stamp[f] = 0;
        id = P.cv_id[slot] = 1;
    }
    for (int d = 0; d < nd; d++) {
        const int r = drow[d];
        const double y = dy[d];
        for (int q = C.rs[r]; q < C.rs[r + 1]; q++) {
            const int f = C.rc[q];
            if (is_forb(C, f)) continue;
            if (stamp[f] != id) {
                stamp[f] = id;
                load[f] = y;
            } else {
                load[f] += y;
            }
        }
    }
    // rows with the fewest free columns first, ties in row order
    int *av = P.cv_av + (size_t)slot * P.cvmax;
    int *ord = P.cv_ord + (size_t)slot * P.cvmax;
    for (int d = 0; d < nd; d++) {
        av[d] = free_in_row(C, drow[d]);
        int j = d;
        while (j > 0 && av[ord[j - 1]] > av[d]) {
            ord[j] = ord[j - 1];
            j--;
        }
        ord[j] = d;
    }
    for (int t = 0; t < nd; t++) {
        const int r = drow[ord[t]];
        double slackv = 1.0;
        for (int q = C.rs[r]; q < C.rs[r + 1]; q++) {
            const int f = C.rc[q];
            if (is_forb(C, f)) continue;
            const double s = 1.0 - load[f];
            if (s < slackv) slackv = s;
        }
        if (slackv <= 1e-12) continue;
        for (int q = C.rs[r]; q < C.rs[r + 1]; q++) {
            const int f = C.rc[q];
            if (!is_forb(C, f)) load[f] += slackv;
        }
        sum += slackv;
        if (sum > lim) {
            apruned++;
            return true;
        }
    }
    return false;
}

// Does column f (joining a path of length a) close kernel set s: |path + f| meets s in exactly floor(|s|/2)?
__device__ bool closes_set(const GmDev &P, const Level &in, uint32_t i, int a, int f, int s) {
    const int b0 = P.ks_start[s], b1 = P.ks_start[s + 1];
    int cnt = 0;
    for (int t = b0; t < b1; t++) {
        const int g = P.ks_cols[t];
        bool in_path = g == f;
        for (int p = 0; p < a && !in_path; p++) in_path = path_at(P, in, i, p) == g;
        cnt += in_path;
    }
    return cnt == (b1 - b0) / 2;
}

// Child of node i (depth a) by column f into slot j of `out`.  pos >= 0: the general branch, whose child forbids
// the pattern-admissible columns of row piv before position pos, f itself, the exclusion partners of f and the
// kernel sets f closes; pos < 0: a leaf-probe hit, which forbids f only (it is completed, never expanded).  Each
// word is assembled in registers and stored once.
__device__ void write_child(const GmDev &P, const Level &in, const Ctx &C, int a, const Level &out, uint32_t j,
                            int f, int how, int piv, int pos, int nwords) {
    const uint32_t i = C.i;
    NODE(out, P.off_hdr, j) = NODE(in, P.off_hdr, i) + (how == 2 ? 1 : 0);
    for (int w = 0; w < P.off_s - P.off_path; w++) {
        uint64_t v = NODE(in, P.off_path + w, i);
        if (w == (a >> 2)) {
            const int sh = (a & 3) * 16;
            v = (v & ~(0xFFFFull << sh)) | ((uint64_t)f << sh);
        }
        NODE(out, P.off_path + w, j) = v;
    }
    for (int w = 0; w < P.W; w++) NODE(out, P.off_s + w, j) = syn(C, w) ^ P.col[(size_t)f * P.W + w];
    for (int t = 0; t < P.ntab; t++) {
        NODE(out, P.off_zh + t, j) = NODE(in, P.off_zh + t, i) ^ P.tzh[t][f];
        NODE(out, P.off_zl + t, j) = NODE(in, P.off_zl + t, i) ^ P.tzl[t][f];
    }
    NODE(out, P.off_lg, j) = NODE(in, P.off_lg, i) ^ P.lg[f];
    const int Nw = P.off_used - P.off_forb;
    for (int w = 0; w < Nw; w++) {
        uint64_t v = forb_word(C, w);
        if ((f >> 6) == w) v |= 1ull << (f & 63);
        if (pos >= 0) {
            for (int q = C.rs[piv]; q < pos; q++) {
                const int g = C.rc[q];
                if ((g >> 6) == w && !is_forb(C, g) && pat_how(P, in, i, g)) v |= 1ull << (g & 63);
            }
            if (P.ex_start)
                for (int q = P.ex_start[f]; q < P.ex_start[f + 1]; q++) {
                    const int g = P.ex_cols[q];
                    if ((g >> 6) == w) v |= 1ull << (g & 63);
                }
            for (int q = P.nks ? P.kc_start[f] : 0; P.nks && q < P.kc_start[f + 1]; q++) {
                const int s = P.kc_ids[q];
                if (!closes_set(P, in, i, a, f, s)) continue;
                for (int t = P.ks_start[s]; t < P.ks_start[s + 1]; t++) {
                    const int g = P.ks_cols[t];
                    if ((g >> 6) == w) v |= 1ull << (g & 63);
                }
            }
        }
        NODE(out, P.off_forb + w, j) = v;
    }
    for (int w = P.off_used; w < nwords; w++) {
        uint64_t v = NODE(in, w, i);
        if (how == 1 && P.cls && (P.cls[f] >> 6) == w - P.off_used) v |= 1ull << (P.cls[f] & 63);
        NODE(out, w, j) = v;
    }
}

// One launch: nodes [first, first + n) of depth a (buffer `in`) are visited; children go to `out` (depth a + 1).
// Dynamic LDS: the incidence arrays when P.inc_lds, then the visited nodes' syndrome and forbidden words (word w of
// thread t at w * kBlock + t) when P.stage.
__global__ __launch_bounds__(kBlock) void k_gm_expand(const GmDev P, const Level in, uint32_t first, uint32_t n,
                                                      int a, const Level out, int nwords, GmFlags *fl,
                                                      unsigned long long *stats) {
    extern __shared__ uint64_t lds_dyn[];
    __shared__ unsigned cnt[S_N];
    __shared__ unsigned nchild_sh[kBlock];
    __shared__ unsigned base_sh;
    const int tid = threadIdx.x;
    if (tid < S_N) cnt[tid] = 0;
    if (blockIdx.x == 0 && tid == 0) atomicSub(&fl->count[a], n);

    const uint32_t t = blockIdx.x * kBlock + tid;
    const uint32_t i = first + t;
    const bool live = t < n;
    Ctx C;
    C.i = i;
    C.rs = P.row_start;
    C.rc = P.row_cols;
    C.cs = P.col_start;
    C.cr = P.col_rows;
    uint64_t *stage = lds_dyn;
    if (P.inc_lds) {
        int *inc = reinterpret_cast<int *>(lds_dyn);
        const int nnz = P.nnz;
        for (int q = tid; q <= P.M; q += kBlock) inc[q] = P.row_start[q];
        for (int q = tid; q < nnz; q += kBlock) inc[P.M + 1 + q] = P.row_cols[q];
        for (int q = tid; q <= P.N; q += kBlock) inc[P.M + 1 + nnz + q] = P.col_start[q];
        for (int q = tid; q < nnz; q += kBlock) inc[P.M + P.N + 2 + nnz + q] = P.col_rows[q];
        C.rs = inc;
        C.rc = inc + P.M + 1;
        C.cs = inc + P.M + 1 + nnz;
        C.cr = inc + P.M + P.N + 2 + nnz;
        stage = lds_dyn + P.inc_words;
    }
    const int Nw = P.off_used - P.off_forb;
    if (P.stage) {
        if (live) {
            for (int w = 0; w < P.W; w++) stage[w * kBlock + tid] = NODE(in, P.off_s + w, i);
            for (int w = 0; w < Nw; w++) stage[(P.W + w) * kBlock + tid] = NODE(in, P.off_forb + w, i);
        }
        C.s = stage + tid;
        C.ss = kBlock;
        C.fb = stage + P.W * kBlock + tid;
        C.fs = kBlock;
    } else {
        C.s = in.base + (size_t)P.off_s * in.cap + i;
        C.ss = in.cap;
        C.fb = in.base + (size_t)P.off_forb * in.cap + i;
        C.fs = in.cap;
    }
    __syncthreads();

    unsigned c_lookups = 0, c_hits = 0, c_false = 0, c_tasks = 0, c_cpr = 0, c_apr = 0, c_cc = 0, c_na = 0,
             c_na1 = 0;
    int piv = -1;          // >= 0: general branch over row piv
    unsigned nchild = 0;
    const int R = P.K - a;
    const bool leaf = P.main_tab >= 0 && R - 1 <= P.tab_k[P.main_tab];

    if (live && !stopped(fl)) {
        c_na = 1;
        int pc = 0;
        for (int w = 0; w < P.W; w++) pc += __popcll(syn(C, w));
        bool go = true;
        if (pc == 0) {
            if (sol_nonzero(P, in, i, a, nullptr, 0, NODE(in, P.off_lg, i)) != (bool)P.kernel_mode)
                record(P, fl, in, i, a, nullptr, 0);
            go = false;
        } else if (R <= 0 || pc > P.wmax * R) {
            go = false;
        } else if (P.cover && pc > R && 4 * pc >= P.cover * R && (P.main_tab < 0 || R > P.tab_k[P.main_tab])) {
            c_cc = 1;
            if (cover_prune(P, C, R, pc, t, c_apr)) {
                c_cpr = 1;
                go = false;
            }
        }
        for (int tb = 0; go && tb < P.ntab; tb++) {
            if (R > P.tab_k[tb]) continue;
            c_lookups++;
            const uint64_t label = P.full_lg ? 1ull : NODE(in, P.off_lg, i);
            const int cost = tab_probe(P.tab[tb], P.tab_wide[tb], NODE(in, P.off_zh + tb, i),
                                       NODE(in, P.off_zl + tb, i), label, R);
            if (cost < 0) {
                go = false;
            } else if (tb == P.main_tab) {
                // the remainder exists up to a fingerprint collision: derive it, with the full budget R
                if (complete(P, in, C, a, R, fl) > 0) c_hits++;
                else c_false++;
                go = false;
            }
        }
        if (go && a == P.td) c_tasks = 1;
        if (go) {
            int best = INT_MAX;
            for (int w = 0; w < P.W; w++) {
                uint64_t x = syn(C, w);
                while (x) {
                    const int r = w * 64 + __builtin_ctzll(x);
                    x &= x - 1;
                    const int av = free_in_row(C, r);
                    if (av < best) {
                        best = av;
                        piv = r;
                    }
                }
            }
            if (best == 0) piv = -1;
        }
        if (piv >= 0 && leaf) {
            // children decided by one probe each: store only the hits
            const int mt = P.main_tab;
            for (int q = C.rs[piv]; q < C.rs[piv + 1]; q++) {
                const int f = C.rc[q];
                if (is_forb(C, f)) continue;
                const int how = pat_how(P, in, i, f);
                if (!how) continue;
                const uint64_t zh = NODE(in, P.off_zh + mt, i) ^ P.tzh[mt][f];
                const uint64_t zl = NODE(in, P.off_zl + mt, i) ^ P.tzl[mt][f];
                const uint64_t label = P.full_lg ? 1ull : (NODE(in, P.off_lg, i) ^ P.lg[f]);
                if (tab_probe(P.tab[mt], P.tab_wide[mt], zh, zl, label, R - 1) < 0) {
                    c_na1++;
                    c_lookups++;
                    continue;
                }
                const unsigned j = atomicAdd(&fl->count[a + 1], 1u);
                if (j >= out.cap) {
                    __hip_atomic_store(&fl->overflow, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    continue;
                }
                write_child(P, in, C, a, out, j, f, how, piv, -1, nwords);
            }
            piv = -1;
        } else if (piv >= 0) {
            for (int q = C.rs[piv]; q < C.rs[piv + 1]; q++) {
                const int f = C.rc[q];
                if (!is_forb(C, f) && pat_how(P, in, i, f)) nchild++;
            }
        }
    }
    // reserve the general branches' children for the whole block at once
    nchild_sh[tid] = nchild;
    __syncthreads();
    if (tid == 0) {
        unsigned run = 0;
        for (int q = 0; q < kBlock; q++) {
            const unsigned c = nchild_sh[q];
            nchild_sh[q] = run;
            run += c;
        }
        base_sh = run ? atomicAdd(&fl->count[a + 1], run) : 0;
    }
    __syncthreads();
    if (nchild) {
        uint32_t j = base_sh + nchild_sh[tid];
        for (int q = C.rs[piv]; q < C.rs[piv + 1]; q++) {
            const int f = C.rc[q];
            if (is_forb(C, f)) continue;
            const int how = pat_how(P, in, i, f);
            if (!how) continue;
            if (j >= out.cap) {
                __hip_atomic_store(&fl->overflow, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            write_child(P, in, C, a, out, j, f, how, piv, q, nwords);
            j++;
        }
    }
    // per wave first, then one LDS atomic per wave and counter
    unsigned c[S_N];
    c[S_LOOKUPS] = c_lookups;
    c[S_HITS] = c_hits;
    c[S_FALSE] = c_false;
    c[S_TASKS] = c_tasks;
    c[S_CPRUNED] = c_cpr;
    c[S_APRUNED] = c_apr;
    c[S_CCALLS] = c_cc;
    c[S_NODES_A] = c_na;
    c[S_NODES_A1] = c_na1;
#pragma unroll
    for (int k = 0; k < S_N; k++) {
        unsigned v = c[k];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) v += (unsigned)__shfl_xor((int)v, o, 64);
        if ((tid & 63) == 0 && v) atomicAdd(&cnt[k], v);
    }
    __syncthreads();
    if (tid < S_N && cnt[tid]) {
        unsigned long long *dst;
        if (tid == S_NODES_A) dst = stats + a * kStatStride;
        else if (tid == S_NODES_A1) dst = stats + (a + 1) * kStatStride;
        else dst = stats + (kStatBase + tid) * kStatStride;
        atomicAdd(dst, (unsigned long long)cnt[tid]);
    }
}

__global__ void k_gm_init(GmFlags *fl, unsigned long long *stats, int nstats) {
    const int t = threadIdx.x;
    if (t < nstats) stats[t * kStatStride] = 0;
    if (t == 0) {
        fl->stop = 0;
        fl->overflow = 0;
        fl->best = INT_MAX;
        fl->nsol = 0;
        fl->nfound = 0;
    }
    if (t < kDmax + 2) fl->count[t] = 0;
}

// process-wide options of the next searches (the reference's gm_set_* globals, cdist/gmitm.c:230-253), per thread
struct Options {
    int cover = 0, ascent = 0;
    int nks = 0;
    const int *ks_start = nullptr, *ks_cols = nullptr, *kc_start = nullptr, *kc_ids = nullptr;
    int kernel_mode = 0;
    const uint64_t *full_lg = nullptr;
    int lw = 0;
    const int *cls = nullptr;
    const uint8_t *onpat = nullptr;
    int nclass = 0, slack = 0;
    int chunk = 0;                 // nodes per launch; 0: the default
};
thread_local Options g_opt;
thread_local int g_last_launches = 0, g_last_chunk = 0;

// Device buffers of the searches, kept per device between calls and grown when a search needs more (the levels'
// node buffers and the dual-ascent scratch run to hundreds of MB; allocating them per call would cost more than a
// small search).  One search at a time per device: the cache's mutex is held for the whole call.
enum {
    B_COL, B_LG, B_FULL, B_RS, B_RC, B_CS, B_CR, B_TZH, B_TZL = B_TZH + kTmax, B_EXS = B_TZL + kTmax, B_EXC, B_KSS,
    B_KSC, B_KCS, B_KCI, B_CLS, B_ONP, B_INV, B_LOAD, B_STAMP, B_CVID, B_DROW, B_DY, B_AV, B_ORD, B_FLAGS, B_STATS,
    B_SOL, B_LEVEL, B_COUNT = B_LEVEL + kDmax + 2
};

struct GmCache {
    std::mutex mu;
    void *ptr[B_COUNT] = {};
    size_t cap[B_COUNT] = {};
    GmFlags *host_flags = nullptr;
    int stamp_n = -1;              // column count of the stamps' slot layout (slot * N + column)
};
constexpr int kMaxDevices = 64;
GmCache g_cache[kMaxDevices];

// at least `bytes` of device memory in slot b; *fresh: newly allocated (contents undefined)
int cache_buf(GmCache &c, int b, size_t bytes, void **out, bool *fresh = nullptr) {
    if (fresh) *fresh = false;
    if (bytes == 0) bytes = 8;
    if (c.cap[b] < bytes) {
        const size_t want = std::max(bytes, c.cap[b] + c.cap[b] / 2);
        if (c.ptr[b]) (void)hipFree(c.ptr[b]);
        c.ptr[b] = nullptr;
        c.cap[b] = 0;
        if (hipMalloc(&c.ptr[b], want) != hipSuccess) {
            c.ptr[b] = nullptr;
            (void)hipGetLastError();
            set_error("device allocation of %zu bytes for the search failed", want);
            return TDP_ENOMEM;
        }
        c.cap[b] = want;
        if (fresh) *fresh = true;
    }
    *out = c.ptr[b];
    return TDP_OK;
}

template <typename T>
int cache_put(GmCache &c, int b, const T *h, size_t n, const T **out) {
    *out = nullptr;
    if (!h || !n) return TDP_OK;
    void *d = nullptr;
    int rc = cache_buf(c, b, n * sizeof(T), &d);
    if (rc != TDP_OK) return rc;
    TDP_HIP_CHECK(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
    *out = static_cast<const T *>(d);
    return TDP_OK;
}

#define GM_TRY(expr)                 \
    do {                             \
        int _rc = (expr);            \
        if (_rc != TDP_OK) return _rc; \
    } while (0)

}  // namespace
}  // namespace tdp

using namespace tdp;

extern "C" {

void tdp_gm_set_cover(int gate4, int ascent) {
    g_opt.cover = gate4;
    g_opt.ascent = ascent;
}

void tdp_gm_set_kernel(int nks, const int *ks_start, const int *ks_cols, const int *kc_start, const int *kc_ids) {
    g_opt.nks = nks;
    g_opt.ks_start = ks_start;
    g_opt.ks_cols = ks_cols;
    g_opt.kc_start = kc_start;
    g_opt.kc_ids = kc_ids;
}

void tdp_gm_set_kernel_mode(int on) { g_opt.kernel_mode = on; }

void tdp_gm_set_full_labels(const uint64_t *lg, int lw) {
    g_opt.full_lg = lg;
    g_opt.lw = lw;
}

void tdp_gm_set_pattern(const int *cls, const uint8_t *onpat, int nclass, int slack) {
    g_opt.cls = cls;
    g_opt.onpat = onpat;
    g_opt.nclass = nclass;
    g_opt.slack = slack;
}

int tdp_gm_num_threads(void) { return 1 << 19; }

int tdp_gm_search(int W, int N, int M, int wmax, const uint64_t *col, const uint64_t *lg, const int *row_start,
                  const int *row_cols, const int *col_start, const int *col_rows, int ntab, void **tabs,
                  const uint64_t **tzh, const uint64_t **tzl, int main_tab, const int *rank, int n_anchor,
                  const int *anchors, const int *ex_start, const int *ex_cols, int K, int enum_all, int task_depth,
                  double time_limit, int64_t *nodes_out, int64_t *stats_out, int *best, int sol_cap, int sol_stride,
                  int *sol_buf, int *nsol, int *status) {
    *status = -1;
    TDP_REQUIRE(best && nsol && nodes_out && stats_out && status, "tdp_gm_search: null output");
    *best = -1;
    *nsol = 0;
    for (int d = 0; d <= kDmax; d++) nodes_out[d] = 0;
    for (int s = 0; s < 8; s++) stats_out[s] = 0;
    if (K > kDmax) {
        *status = 4;
        return TDP_OK;
    }
    const Options opt = g_opt;
    g_last_launches = 0;
    TDP_REQUIRE(W >= 1 && N >= 1 && M >= 1 && M <= 64 * W && N < 65536, "tdp_gm_search: W %d N %d M %d", W, N, M);
    TDP_REQUIRE(ntab >= 0 && ntab <= kTmax && main_tab < ntab && K >= 0 && n_anchor >= 0,
                "tdp_gm_search: %d tables (main %d), K %d", ntab, main_tab, K);
    TDP_REQUIRE(sol_cap >= 0 && (sol_cap == 0 || (sol_buf && sol_stride >= 1)), "tdp_gm_search: solution buffer");
    TDP_REQUIRE(rank && (n_anchor == 0 || anchors), "tdp_gm_search: rank and anchors are required");
    const auto t0 = std::chrono::steady_clock::now();
    const int Nw = (N + 63) / 64;
    const int Cw = opt.cls ? std::max(1, (opt.nclass + 63) / 64) : 0;
    const int KL = std::max(K, 1);           // deepest level held: the anchors are visited even at K = 0
    const int PW = (KL + 3) / 4;
    GmDev P{};
    P.W = W;
    P.N = N;
    P.M = M;
    P.wmax = wmax;
    P.K = K;
    P.td = std::min(std::max(task_depth, 1), std::max(K, 1));
    P.enum_all = enum_all;
    P.kernel_mode = opt.kernel_mode;
    P.lw = opt.full_lg ? opt.lw : 0;
    P.cover = wmax > kInvMax ? 0 : opt.cover;
    P.ascent = opt.ascent;
    P.ntab = ntab;
    P.main_tab = main_tab;
    P.off_hdr = 0;
    P.off_path = 1;
    P.off_s = P.off_path + PW;
    P.off_zh = P.off_s + W;
    P.off_zl = P.off_zh + ntab;
    P.off_lg = P.off_zl + ntab;
    P.off_forb = P.off_lg + 1;
    P.off_used = P.off_forb + Nw;
    const int nwords = P.off_used + Cw;
    P.inv_wmax = 1.0 / wmax;
    P.nks = opt.nks;
    P.slack = opt.slack;
    P.sol_cap = sol_cap;
    P.sol_stride = sol_stride;

    int dev = 0;
    TDP_HIP_CHECK(hipGetDevice(&dev));
    TDP_REQUIRE(dev >= 0 && dev < kMaxDevices, "tdp_gm_search: device %d beyond %d", dev, kMaxDevices);
    GmCache &G = g_cache[dev];
    std::lock_guard<std::mutex> hold(G.mu);

    const size_t nnz = (size_t)row_start[M];
    TDP_REQUIRE((size_t)col_start[N] == nnz, "tdp_gm_search: row and column incidence disagree");
    P.nnz = (int)nnz;
    GM_TRY(cache_put(G, B_COL, col, (size_t)N * W, &P.col));
    std::vector<uint64_t> zero_lg;
    if (!lg) zero_lg.assign(N, 0);
    GM_TRY(cache_put(G, B_LG, lg ? lg : zero_lg.data(), (size_t)N, &P.lg));
    if (opt.full_lg) GM_TRY(cache_put(G, B_FULL, opt.full_lg, (size_t)N * opt.lw, &P.full_lg));
    const int one = 0;
    GM_TRY(cache_put(G, B_RS, row_start, (size_t)M + 1, &P.row_start));
    GM_TRY(cache_put(G, B_RC, nnz ? row_cols : &one, nnz ? nnz : 1, &P.row_cols));
    GM_TRY(cache_put(G, B_CS, col_start, (size_t)N + 1, &P.col_start));
    GM_TRY(cache_put(G, B_CR, nnz ? col_rows : &one, nnz ? nnz : 1, &P.col_rows));
    for (int t = 0; t < ntab; t++) {
        DevTable *tb = static_cast<DevTable *>(tabs[t]);
        TDP_REQUIRE(tb != nullptr, "tdp_gm_search: table %d is null", t);
        P.tab[t] = view(tb);
        P.tab_wide[t] = tb->wide;
        P.tab_k[t] = tb->k;
        TDP_REQUIRE(t != main_tab || tb->k <= kCompleteMax, "tdp_gm_search: exact table depth %d above %d", tb->k,
                    kCompleteMax);
        GM_TRY(cache_put(G, B_TZH + t, tzh[t], (size_t)N, &P.tzh[t]));
        GM_TRY(cache_put(G, B_TZL + t, tzl[t], (size_t)N, &P.tzl[t]));
    }
    if (ex_start) {
        GM_TRY(cache_put(G, B_EXS, ex_start, (size_t)N + 1, &P.ex_start));
        GM_TRY(cache_put(G, B_EXC, ex_cols, (size_t)std::max(ex_start[N], 1), &P.ex_cols));
    }
    if (opt.nks > 0) {
        GM_TRY(cache_put(G, B_KSS, opt.ks_start, (size_t)opt.nks + 1, &P.ks_start));
        GM_TRY(cache_put(G, B_KSC, opt.ks_cols, (size_t)std::max(opt.ks_start[opt.nks], 1), &P.ks_cols));
        GM_TRY(cache_put(G, B_KCS, opt.kc_start, (size_t)N + 1, &P.kc_start));
        GM_TRY(cache_put(G, B_KCI, opt.kc_ids, (size_t)std::max(opt.kc_start[N], 1), &P.kc_ids));
    }
    if (opt.cls) {
        for (int j = 0; j < N; j++)
            TDP_REQUIRE(opt.cls[j] >= -1 && opt.cls[j] < opt.nclass, "tdp_gm_search: class %d of column %d",
                        opt.cls[j], j);
        GM_TRY(cache_put(G, B_CLS, opt.cls, (size_t)N, &P.cls));
        std::vector<uint8_t> on(std::max(opt.nclass, 1), 0);
        for (int c = 0; c < opt.nclass; c++) on[c] = opt.onpat[c];
        GM_TRY(cache_put(G, B_ONP, on.data(), on.size(), &P.onpat));
    }
    std::vector<double> inv(kInvMax + 1, 0.0);
    for (int m = 1; m <= kInvMax; m++) inv[m] = 1.0 / m;
    GM_TRY(cache_put(G, B_INV, inv.data(), inv.size(), &P.inv));

    // LDS: the incidence arrays and the visited nodes' syndrome and forbidden words, when they fit
    const size_t inc_ints = (size_t)M + N + 2 + 2 * nnz;
    P.inc_lds = inc_ints * 4 <= (size_t)kIncLdsBytes;
    P.inc_words = P.inc_lds ? (int)((inc_ints + 1) / 2) : 0;
    P.stage = W + Nw <= kStageWords;
    const size_t lds_bytes = ((size_t)P.inc_words + (P.stage ? (size_t)(W + Nw) * kBlock : 0)) * 8;

    int deg = 1;
    for (int r = 0; r < M; r++) deg = std::max(deg, row_start[r + 1] - row_start[r]);
    // chunk: nodes per launch.  Buffers: depth 1 holds the anchors, deeper depths chunk x degree nodes each.
    // A node's visit is a long chain of dependent loads (tens of microseconds), so a launch takes about that long
    // whatever its size until the device is full: the default chunk is about the device's resident threads.
    uint32_t chunk = opt.chunk > 0 ? (uint32_t)std::max(opt.chunk, kBlock) : (1u << 19);
    const size_t node_bytes = (size_t)nwords * 8;
    const size_t level_budget = (size_t)32 << 30;
    while (chunk > (uint32_t)kBlock && (size_t)KL * chunk * deg * node_bytes > level_budget) chunk /= 2;
    P.cvmax = std::min(kCvMax, std::max(1, wmax * KL));
    if (P.cover && P.ascent) {
        const size_t per_slot = (size_t)N * 12 + (size_t)P.cvmax * 20 + 4;
        while (chunk > (uint32_t)kBlock && chunk * per_slot > ((size_t)8 << 30)) chunk /= 2;
        void *p = nullptr;
        bool f1 = false, f2 = false;
        GM_TRY(cache_buf(G, B_LOAD, (size_t)chunk * N * 8, &p));
        P.cv_load = static_cast<double *>(p);
        GM_TRY(cache_buf(G, B_STAMP, (size_t)chunk * N * 4, &p, &f1));
        P.cv_stamp = static_cast<unsigned *>(p);
        GM_TRY(cache_buf(G, B_CVID, (size_t)chunk * 4, &p, &f2));
        P.cv_id = static_cast<unsigned *>(p);
        // a slot's stamps are compared with that slot's own running id: they start together at zero, and again
        // whenever the slot layout changes with N (a region would otherwise hold another slot's ids)
        if (f1 || f2 || G.stamp_n != N) {
            TDP_HIP_CHECK(hipMemset(G.ptr[B_STAMP], 0, G.cap[B_STAMP]));
            TDP_HIP_CHECK(hipMemset(G.ptr[B_CVID], 0, G.cap[B_CVID]));
            G.stamp_n = N;
        }
        GM_TRY(cache_buf(G, B_DROW, (size_t)chunk * P.cvmax * 4, &p));
        P.cv_drow = static_cast<int *>(p);
        GM_TRY(cache_buf(G, B_DY, (size_t)chunk * P.cvmax * 8, &p));
        P.cv_dy = static_cast<double *>(p);
        GM_TRY(cache_buf(G, B_AV, (size_t)chunk * P.cvmax * 4, &p));
        P.cv_av = static_cast<int *>(p);
        GM_TRY(cache_buf(G, B_ORD, (size_t)chunk * P.cvmax * 4, &p));
        P.cv_ord = static_cast<int *>(p);
    }
    void *pv = nullptr;
    GM_TRY(cache_buf(G, B_FLAGS, sizeof(GmFlags), &pv));
    GmFlags *d_fl = static_cast<GmFlags *>(pv);
    const int nstats = kStatBase + S_N;
    GM_TRY(cache_buf(G, B_STATS, nstats * 8 * kStatStride, &pv));
    unsigned long long *d_stats = static_cast<unsigned long long *>(pv);
    if (sol_cap > 0) {
        GM_TRY(cache_buf(G, B_SOL, (size_t)sol_cap * sol_stride * 4, &pv));
        P.sol = static_cast<int *>(pv);
    }
    if (!G.host_flags)
        TDP_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&G.host_flags), sizeof(GmFlags), hipHostMallocDefault));
    GmFlags *h_fl = G.host_flags;
    k_gm_init<<<1, 128>>>(d_fl, d_stats, nstats);
    TDP_LAUNCH_CHECK();

    // depth-1 nodes: the anchors, each forbidding the columns of lower rank, itself, its exclusion partners and
    // the kernel sets it closes (cdist/gmitm.c:727-740, 744-760)
    std::vector<Level> lev(kDmax + 2, Level{nullptr, 0});
    std::vector<uint32_t> count(kDmax + 2, 0);
    if (n_anchor > 0) {
        const uint32_t cap = (uint32_t)n_anchor;
        std::vector<uint64_t> h((size_t)nwords * cap, 0);
        auto at = [&](int w, uint32_t i) -> uint64_t & { return h[(size_t)w * cap + i]; };
        for (uint32_t i = 0; i < cap; i++) {
            const int f = anchors[i];
            TDP_REQUIRE(f >= 0 && f < N, "tdp_gm_search: anchor %d out of range", f);
            at(P.off_path, i) = (uint64_t)f;
            for (int w = 0; w < W; w++) at(P.off_s + w, i) = col[(size_t)f * W + w];
            for (int t = 0; t < ntab; t++) {
                at(P.off_zh + t, i) = tzh[t][f];
                at(P.off_zl + t, i) = tzl[t][f];
            }
            at(P.off_lg, i) = lg ? lg[f] : 0;
            auto forbid = [&](int g) { at(P.off_forb + (g >> 6), i) |= 1ull << (g & 63); };
            for (int g = 0; g < N; g++)
                if (rank[g] < rank[f]) forbid(g);
            forbid(f);
            if (ex_start)
                for (int q = ex_start[f]; q < ex_start[f + 1]; q++) forbid(ex_cols[q]);
            for (int q = opt.nks ? opt.kc_start[f] : 0; opt.nks && q < opt.kc_start[f + 1]; q++) {
                const int s = opt.kc_ids[q];
                if ((opt.ks_start[s + 1] - opt.ks_start[s]) / 2 == 1)   // the anchor alone reaches half the set
                    for (int t = opt.ks_start[s]; t < opt.ks_start[s + 1]; t++) forbid(opt.ks_cols[t]);
            }
            if (opt.cls) {
                const int cl = opt.cls[f];
                if (cl >= 0 && opt.onpat[cl]) at(P.off_used + (cl >> 6), i) |= 1ull << (cl & 63);
                else if (opt.slack > 0) at(P.off_hdr, i) = 1;
            }
        }
        void *p = nullptr;
        GM_TRY(cache_buf(G, B_LEVEL + 1, h.size() * 8, &p));
        TDP_HIP_CHECK(hipMemcpy(p, h.data(), h.size() * 8, hipMemcpyHostToDevice));
        lev[1] = Level{static_cast<uint64_t *>(p), cap};
        count[1] = cap;
        TDP_HIP_CHECK(hipMemcpy(&d_fl->count[1], &count[1], 4, hipMemcpyHostToDevice));
    }
    const uint32_t deep_cap = chunk * (uint32_t)deg;
    auto level = [&](int d, Level *out) -> int {
        if (d > KL) {
            *out = Level{nullptr, 0};
            return TDP_OK;
        }
        if (!lev[d].base) {
            void *p = nullptr;
            GM_TRY(cache_buf(G, B_LEVEL + d, (size_t)deep_cap * node_bytes, &p));
            lev[d] = Level{static_cast<uint64_t *>(p), deep_cap};
        }
        *out = lev[d];
        return TDP_OK;
    };

    bool timed_out = false, overflow = false, stop = false;
    const double deadline = time_limit > 0 ? time_limit : 0.0;
    while (true) {
        int d = 0;
        for (int q = KL; q >= 1; q--)
            if (count[q]) {
                d = q;
                break;
            }
        if (!d) break;
        if (deadline > 0 &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > deadline) {
            timed_out = true;
            break;
        }
        int src = d;
        uint32_t n = std::min(count[d], chunk);
        if (d >= 2 && count[d] < chunk / 2 && count[d - 1] > 0 && lev[d].cap > count[d]) {
            const uint32_t room = (lev[d].cap - count[d]) / (uint32_t)deg;
            if (room >= 1) {
                src = d - 1;
                n = std::min(std::min(count[d - 1], room), chunk);
            }
        }
        Level out{};
        GM_TRY(level(src + 1, &out));
        const uint32_t first = count[src] - n;
        k_gm_expand<<<(n + kBlock - 1) / kBlock, kBlock, lds_bytes>>>(P, lev[src], first, n, src, out, nwords, d_fl,
                                                                      d_stats);
        TDP_LAUNCH_CHECK();
        TDP_HIP_CHECK(hipMemcpyAsync(h_fl, d_fl, sizeof(GmFlags), hipMemcpyDeviceToHost, 0));
        TDP_HIP_CHECK(hipStreamSynchronize(0));
        g_last_launches++;
        if (h_fl->overflow) {
            overflow = true;
            break;
        }
        for (int q = 1; q <= KL; q++) count[q] = h_fl->count[q];
        if (h_fl->stop) {
            stop = true;
            break;
        }
    }
    TDP_HIP_CHECK(hipMemcpy(h_fl, d_fl, sizeof(GmFlags), hipMemcpyDeviceToHost));
    std::vector<unsigned long long> raw((size_t)nstats * kStatStride);
    TDP_HIP_CHECK(hipMemcpy(raw.data(), d_stats, raw.size() * 8, hipMemcpyDeviceToHost));
    auto st = [&](int k) { return (int64_t)raw[(size_t)k * kStatStride]; };
    TDP_REQUIRE(!overflow, "tdp_gm_search: a node buffer overflowed");
    for (int q = 0; q <= kDmax; q++) nodes_out[q] = st(q);
    stats_out[0] = st(kStatBase + S_LOOKUPS);
    stats_out[1] = st(kStatBase + S_HITS);
    stats_out[2] = st(kStatBase + S_FALSE);
    stats_out[3] = st(kStatBase + S_TASKS);
    stats_out[4] = st(kStatBase + S_CPRUNED);
    stats_out[5] = st(kStatBase + S_APRUNED);
    stats_out[6] = st(kStatBase + S_CCALLS);
    stats_out[7] = (int64_t)h_fl->nfound;
    *best = h_fl->best == INT_MAX ? -1 : h_fl->best;
    *nsol = (int)std::min<unsigned>(h_fl->nsol, (unsigned)sol_cap);
    if (*nsol > 0)
        TDP_HIP_CHECK(hipMemcpy(sol_buf, P.sol, (size_t)*nsol * sol_stride * 4, hipMemcpyDeviceToHost));
    stop = stop || h_fl->stop;
    *status = timed_out ? 2 : (stop ? 1 : 0);
    g_last_chunk = (int)chunk;
    return TDP_OK;
}

void tdp_gm_set_chunk(int nodes) { g_opt.chunk = nodes; }

int tdp_gm_last_launches(void) { return g_last_launches; }

void tdp_gm_release(void) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return;
    GmCache &G = g_cache[dev];
    std::lock_guard<std::mutex> hold(G.mu);
    for (int b = 0; b < B_COUNT; b++) {
        if (G.ptr[b]) (void)hipFree(G.ptr[b]);
        G.ptr[b] = nullptr;
        G.cap[b] = 0;
    }
    if (G.host_flags) (void)hipHostFree(G.host_flags);
    G.host_flags = nullptr;
    G.stamp_n = -1;
}

}  // extern "C"
