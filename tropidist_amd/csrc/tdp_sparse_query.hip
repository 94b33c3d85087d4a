// One sparse query run natively end to end on the GPU (the query of cdist/sparse_dp.py:985-1095 and its sweeps,
// cdist/sparse_dp.py:621-820, with the column step and join of cdist/sparse_dp.c:116-236).
//
// The host side is a small runtime that owns the query's device tables and walks the cuts:
//   backward sweep   from {0: 0} at cut N down to cut 0, keeping weights <= K - a(c) at cut c;
//   forward sweep    from {0: 0} at cut 0 up to cut N, keeping weights <= a(c) - 1 and joining its new or improved
//                    states (weights <= a(c)) with the backward table at every cut; where a(c) drops, the whole
//                    table before the cut is joined too, then cut down to the new bound;
//   traceback        from the best join back to cut 0 through the forward tables and on to cut N through the
//                    backward ones, one device thread per stretch of tables.
// mode "forward" is one sweep keeping weights <= K whose value is state 0 at cut N.  A profile sweep keeps weights
// <= cap in one direction and histograms every table's values (the probes of the split heuristics).
//
// Every column step is one launch over the old table that also builds the new table's hash index as it appends
// (the index, cleared by a fill launch just before, has room for the most outputs the step can make), so no table is
// re-read to be indexed.  The step's counts and the join's least total come back in one read per cut.  Tables are
// kept in full while the kept bytes stay within the caller's budget, and otherwise only at checkpoints every
// `every` cuts; a table that was not kept is recomputed from the checkpoint before it when the joins or the
// traceback need it (a stretch at a time).  Device memory comes from the stream-ordered allocator.
#include <algorithm>
#include <chrono>
#include <vector>

#include "tdp_sparse.h"

namespace tdp {

// ---------------------------------------------------------------------------------------------- kernels
// The step of tdp_sparse.hip's k_sp_step with the new table's index built at append time: every table entry an old
// entry emits is written at its reserved position p and p is placed in the first empty slot of its probe sequence
// (output keys are distinct, so no two entries compete for a key).  counters[0] = table count | changed count << 32,
// counters[2] = overflow flag.
__global__ __launch_bounds__(kSpBlock) void k_sq_step(const Ent *__restrict__ old, uint64_t n,
                                                      const uint32_t *__restrict__ slots, uint64_t capmask,
                                                      StepArgs a, BoundArgs b, Ent *__restrict__ out,
                                                      uint32_t *__restrict__ nslots, uint64_t ncapmask,
                                                      Ent *__restrict__ ch, unsigned long long *counters) {
    sp_step_tile<true>(old, n, (uint64_t)blockIdx.x * kSpBlock * kSpTile, slots, capmask, a, b, out, ch,
                       &counters[0], &counters[2], nslots, ncapmask);
}

__global__ __launch_bounds__(kSpBlock) void k_sq_clear(uint32_t *slots, uint64_t cap) {
    for (uint64_t i = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * kSpBlock)
        slots[i] = kEmpty;
}

// The start table {0: 0} and its index; counters cleared.
__global__ void k_sq_start(Ent *ents, uint32_t *slots, uint64_t capmask, unsigned long long *counters) {
    if (threadIdx.x == 0) {
        ents[0] = Ent{0, 0};
        slots[hash_key(0, 0) & capmask] = 0;
    }
    if (threadIdx.x < 4) counters[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kSpBlock) void k_sq_index(const Ent *__restrict__ ents, uint64_t n, uint32_t *slots,
                                                       uint64_t capmask) {
    uint64_t i = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x;
    if (i < n) sp_insert(ents, i, slots, capmask);
}

__global__ __launch_bounds__(kSpBlock) void k_sq_hist(const Ent *__restrict__ ents, uint64_t n, int top,
                                                      unsigned long long *hist) {
    __shared__ uint32_t bins[256];
    bins[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kSpBlock)
        atomicAdd(&bins[ents[i].hv >> 56], 1u);
    __syncthreads();
    if ((int)threadIdx.x <= top && bins[threadIdx.x])
        atomicAdd(&hist[threadIdx.x], (unsigned long long)bins[threadIdx.x]);
}

__global__ void k_sq_zero(unsigned long long *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

struct SqRef {
    const Ent *ents;
    const uint32_t *slots;
    uint64_t capmask;
    uint64_t pad;
};

// One thread walks a stretch of steps back along a path (the undo rule of cdist/sparse_dp.py:812-819): before step
// p the state was s | bits; it stayed if the table before the step holds it with value v, else it came from
// (s | bits) ^ col with value v - 1 and column p is used.  tabs[k] is the table before step first + k * dir (first,
// first + dir, ... toward the start of the sweep); io: {lo, hi, v, ok}, used positions appended at used[*nused].
__global__ void k_sq_walk(const SqRef *__restrict__ tabs, const uint64_t *__restrict__ prm, int first, int count,
                          int dir, int bits_word, unsigned long long *io, int32_t *used, int32_t *nused, int cap) {
    if (threadIdx.x != 0) return;
    uint64_t slo = io[0], shi = io[1];
    int v = (int)io[2];
    bool ok = io[3] != 0;
    int nu = *nused;
    for (int k = 0; k < count && ok; k++) {
        const int p = first + k * dir;
        const uint64_t *q = prm + (uint64_t)p * 10;
        const uint64_t mlo = slo | q[bits_word], mhi = shi | q[bits_word + 1];
        const SqRef &t = tabs[k];
        if ((int)sp_find(t.ents, t.slots, t.capmask, mlo, mhi) == v) {
            slo = mlo;
            shi = mhi;
            continue;
        }
        if (!(q[0] | q[1]) || v < 1 ||
            (int)sp_find(t.ents, t.slots, t.capmask, mlo ^ q[0], mhi ^ q[1]) != v - 1) {
            ok = false;
            break;
        }
        slo = mlo ^ q[0];
        shi = mhi ^ q[1];
        v -= 1;
        if (nu < cap) used[nu] = p;
        nu++;
    }
    io[0] = slo;
    io[1] = shi;
    io[2] = (unsigned long long)v;
    io[3] = ok ? 1 : 0;
    *nused = nu;
}

namespace {

using Clock = std::chrono::steady_clock;

uint64_t up256(uint64_t x) { return (x + 255) / 256 * 256; }
uint64_t index_cap(uint64_t n) {
    uint64_t c = 64;
    while (c < 2 * n) c <<= 1;
    return c;
}
unsigned grid_for(uint64_t items, uint64_t limit = 1ull << 20) {
    uint64_t g = (items + kSpBlock - 1) / kSpBlock;
    return (unsigned)std::max<uint64_t>(1, std::min(g, limit));
}

// A device table: entries, index, and the allocation both live in (freed when the table is dropped).
struct Tab {
    Ent *ents = nullptr;
    uint64_t n = 0;
    uint32_t *slots = nullptr;
    uint64_t cap = 0;        // index slots
    void *mem = nullptr;
    uint64_t bytes = 0;
};

}  // namespace

struct SqRun {
    const tdp_sq_spec_t &S;
    hipStream_t st;
    Clock::time_point t0;
    unsigned long long *d_cnt = nullptr;      // [0..3] step counters, [4..7] join best, [8..263] histogram
    uint64_t *d_prm = nullptr;                // device copies of the column words and the lookahead classes
    uint64_t *d_bmask[2] = {nullptr, nullptr};
    int64_t *d_bw[2] = {nullptr, nullptr};
    unsigned long long *h_cnt = nullptr;      // pinned mirror
    uint64_t held = 0, peak_held = 0;
    int status = TDP_SQ_DONE;
    int fail_cut = -1;
    uint64_t steps = 0, peak_states = 1;

    explicit SqRun(const tdp_sq_spec_t &s, hipStream_t stream) : S(s), st(stream), t0(Clock::now()) {}

    // Device memory of the query: from the caller's allocator when it gave one (a buffer goes back to it only once
    // the stream is idle), else stream-ordered from the device's default pool.
    int dalloc(void **p, uint64_t bytes) {
        *p = nullptr;
        if (S.dev_alloc) {
            *p = S.dev_alloc(bytes, S.alloc_ctx);
            if (!*p) {
                set_error("sparse query: the caller's allocator has no %llu bytes", (unsigned long long)bytes);
                return TDP_ENOMEM;
            }
            return TDP_OK;
        }
        TDP_HIP_CHECK(hipMallocAsync(p, bytes, st));
        return TDP_OK;
    }
    void dfree(void *p) {
        if (!p) return;
        if (S.dev_free) {
            (void)hipStreamSynchronize(st);
            S.dev_free(p, S.alloc_ctx);
        } else {
            (void)hipFreeAsync(p, st);
        }
    }
    int alloc(Tab &t, uint64_t nents, uint64_t cap) {
        t.cap = cap;
        const uint64_t eb = up256(16 * std::max<uint64_t>(nents, 1));
        t.bytes = eb + 4 * cap;
        int rc = dalloc(&t.mem, t.bytes);
        if (rc) return rc;
        t.ents = static_cast<Ent *>(t.mem);
        t.slots = reinterpret_cast<uint32_t *>(static_cast<uint8_t *>(t.mem) + eb);
        held += t.bytes;
        peak_held = std::max(peak_held, held);
        return TDP_OK;
    }
    void drop(Tab &t) {
        if (t.mem) {
            dfree(t.mem);
            held -= t.bytes;
        }
        t = Tab();
    }
    double elapsed() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    bool past_deadline() const { return S.max_seconds >= 0 && elapsed() > S.max_seconds; }

    int start_table(Tab &t) {
        int rc = alloc(t, 1, 64);
        if (rc) return rc;
        k_sq_clear<<<1, kSpBlock, 0, st>>>(t.slots, t.cap);
        k_sq_start<<<1, 64, 0, st>>>(t.ents, t.slots, t.cap - 1, d_cnt);
        TDP_LAUNCH_CHECK();
        t.n = 1;
        return TDP_OK;
    }

    // The column step at position p of direction dir (+1: cut p -> p + 1, -1: cut p + 1 -> p) with the caps given;
    // want: a changed list (capacity 2n) is written to *ch.  Counts land in d_cnt[0..2] (read by the caller).
    int step(const Tab &old, int p, int dir, int cap, int kb, bool use_h, Tab &out, Ent *ch) {
        const uint64_t *q = S.prm + (uint64_t)p * 10;
        const int cut = dir > 0 ? p + 1 : p;
        const uint64_t n = old.n, ncap = index_cap(2 * std::max<uint64_t>(n, 1));
        if (n >= (1ull << 31)) {
            set_error("sparse query: a table of %llu states (fewer than 2^31 are stepped)", (unsigned long long)n);
            return TDP_EINVAL;
        }
        int rc = alloc(out, 2 * n, ncap);
        if (rc) return rc;
        StepArgs a;
        a.upd_lo = q[0];
        a.upd_hi = q[1] & kKeyMask;
        a.sm_lo = dir > 0 ? q[2] : q[6];
        a.sm_hi = (dir > 0 ? q[3] : q[7]) & kKeyMask;
        a.sb_lo = dir > 0 ? q[4] : q[8];
        a.sb_hi = (dir > 0 ? q[5] : q[9]) & kKeyMask;
        a.cap = cap;
        a.kb = std::min(cap + 1, kb);
        a.want_changed = ch != nullptr;
        a.out_cap = 2 * n;
        a.ch_cap = 2 * n;
        BoundArgs b;
        const int d = dir > 0 ? 0 : 1;
        b.on = use_h ? 1 : 0;
        b.ncls = use_h ? S.bnd_ncls[d][cut] : 0;
        b.mask = use_h ? d_bmask[d] + 2 * S.bnd_off[d][cut] : nullptr;
        b.w = use_h ? d_bw[d] + S.bnd_off[d][cut] : nullptr;
        b.extra = use_h ? S.bnd_extra[d][cut] : 0;
        b.tm_lo = S.tm[2 * cut];
        b.tm_hi = S.tm[2 * cut + 1] & kKeyMask;
        b.K = S.K;
        k_sq_zero<<<1, 64, 0, st>>>(d_cnt, 3);
        k_sq_clear<<<grid_for(ncap, 4096), kSpBlock, 0, st>>>(out.slots, ncap);
        if (n)
            k_sq_step<<<grid_for((n + kSpTile - 1) / kSpTile, ~0ull), kSpBlock, 0, st>>>(
                old.ents, n, old.slots, old.cap - 1, a, b, out.ents, out.slots, ncap - 1, ch, d_cnt);
        TDP_LAUNCH_CHECK();
        steps++;
        return TDP_OK;
    }

    // Reads d_cnt[0..nwords) after the stream's work.
    int read(int nwords) {
        TDP_HIP_CHECK(hipMemcpyAsync(h_cnt, d_cnt, 8ull * nwords, hipMemcpyDeviceToHost, st));
        TDP_HIP_CHECK(hipStreamSynchronize(st));
        return TDP_OK;
    }

    // Reads a step's counters: h_cnt[0] table count, [1] changed count, [2] overflow flag.
    int read_step() {
        int rc = read(3);
        if (rc) return rc;
        h_cnt[1] = h_cnt[0] >> 32;
        h_cnt[0] &= 0xFFFFFFFFull;
        return TDP_OK;
    }

    int check_budget(uint64_t n, uint64_t nc, int cut) {
        peak_states = std::max(peak_states, std::max(n, nc));
        if (S.max_states && std::max(n, nc) > S.max_states) {
            status = TDP_SQ_STATES;
            fail_cut = cut;
            return 1;
        }
        if (past_deadline()) {
            status = TDP_SQ_DEADLINE;
            fail_cut = cut;
            return 1;
        }
        return 0;
    }
};

namespace {

// A sweep's kept tables: every cut while the kept bytes stay within the budget, afterwards only the checkpoints.
struct Kept {
    std::vector<Tab> at;     // per cut
    uint64_t bytes = 0;
};

bool is_checkpoint(int c, int N, int every) { return c % every == 0 || c == 0 || c == N; }

}  // namespace

// Sweep of one direction without frontier (backward half, forward mode, profiles): tables at every cut from the
// start; caps[c] = the weights kept at cut c; hist (optional): per cut the counts of values 0..top.
static int plain_sweep(SqRun &R, int dir, const int32_t *caps, bool use_h, Kept &kept, uint64_t *counts,
                       uint64_t *hist, int top) {
    const tdp_sq_spec_t &S = R.S;
    const int N = S.N;
    kept.at.assign(N + 1, Tab());
    kept.bytes = 0;
    int c = dir > 0 ? 0 : N;
    Tab cur;
    int rc = R.start_table(cur);
    if (rc) return rc;
    if (counts) counts[c] = 1;
    if (hist) hist[(uint64_t)c * (top + 1)] = 1;
    kept.at[c] = cur;
    kept.bytes += cur.bytes;
    bool cur_kept = true;
    for (int k = 0; k < N; k++) {
        const int p = dir > 0 ? k : N - 1 - k;
        const int nc = dir > 0 ? p + 1 : p;
        Tab nxt;
        rc = R.step(cur, p, dir, caps[nc], 1 << 30, use_h, nxt, nullptr);
        if (rc) return rc;
        rc = R.read_step();
        if (rc) return rc;
        if (R.h_cnt[2]) {
            set_error("sparse query: step output past its capacity");
            return TDP_EHIP;
        }
        nxt.n = R.h_cnt[0];
        if (hist) {
            k_sq_zero<<<1, 256, 0, R.st>>>(R.d_cnt + 8, 256);
            k_sq_hist<<<grid_for(nxt.n, 256), kSpBlock, 0, R.st>>>(nxt.ents, nxt.n, top, R.d_cnt + 8);
            TDP_LAUNCH_CHECK();
            TDP_HIP_CHECK(hipMemcpyAsync(R.h_cnt + 8, R.d_cnt + 8, 8ull * (top + 1), hipMemcpyDeviceToHost, R.st));
            TDP_HIP_CHECK(hipStreamSynchronize(R.st));
            for (int w = 0; w <= top; w++) hist[(uint64_t)nc * (top + 1) + w] = R.h_cnt[8 + w];
        }
        if (counts) counts[nc] = nxt.n;
        if (!cur_kept) R.drop(cur);
        cur = nxt;
        cur_kept = is_checkpoint(nc, N, S.every) || kept.bytes + nxt.bytes <= S.keep_bytes;
        if (cur_kept) {
            kept.at[nc] = nxt;
            kept.bytes += nxt.bytes;
        }
        if (R.check_budget(nxt.n, 0, nc)) {
            if (!cur_kept) R.drop(cur);
            return TDP_OK;
        }
    }
    if (!cur_kept) R.drop(cur);
    return TDP_OK;
}

// The table of a sweep at cut c, recomputing the stretch from the nearest kept table before it (in the sweep's
// direction) into `cache` (cut -> table, replaced per stretch).  Non-frontier sweeps only.
static int table_at(SqRun &R, int dir, const int32_t *caps, bool use_h, Kept &kept, std::vector<Tab> &cache,
                    int c, Tab &out) {
    const int N = R.S.N;
    if (kept.at[c].mem) {
        out = kept.at[c];
        return TDP_OK;
    }
    if ((int)cache.size() == N + 1 && cache[c].mem) {
        out = cache[c];
        return TDP_OK;
    }
    for (Tab &t : cache) R.drop(t);
    cache.assign(N + 1, Tab());
    int from = c;
    while (!kept.at[from].mem) from -= dir;                // the kept table before c in the sweep's direction
    // recompute up to the next kept table (or the end of the sweep)
    Tab cur = kept.at[from];
    for (int x = from; x != c || true;) {
        const int p = dir > 0 ? x : x - 1;
        const int nx = x + dir;
        if (nx < 0 || nx > N || kept.at[nx].mem) break;
        Tab nxt;
        int rc = R.step(cur, p, dir, caps[nx], 1 << 30, use_h, nxt, nullptr);
        if (rc) return rc;
        rc = R.read_step();
        if (rc) return rc;
        nxt.n = R.h_cnt[0];
        cache[nx] = nxt;
        cur = nxt;
        x = nx;
    }
    out = cache[c];
    if (!out.mem) {
        set_error("sparse query: table at cut %d was not recomputed", c);
        return TDP_EINVAL;
    }
    return TDP_OK;
}

// Walks back along the path from (state, value) at cut `cut` to the start of a non-frontier sweep of direction dir,
// appending used positions.
static int walk_back(SqRun &R, int dir, const int32_t *caps, bool use_h, Kept &kept, std::vector<Tab> &cache,
                     int cut, uint64_t slo, uint64_t shi, int v, std::vector<int32_t> &used, bool &ok) {
    const tdp_sq_spec_t &S = R.S;
    const int N = S.N;
    const int end = dir > 0 ? 0 : N;
    const int bits_word = dir > 0 ? 4 : 8;
    std::vector<SqRef> refs;
    int c = cut;
    void *dev = nullptr;
    const uint64_t dev_bytes = up256(sizeof(SqRef) * (N + 1)) + 64 + up256(4ull * (S.K + 2) + 8);
    {
        int rc0 = R.dalloc(&dev, dev_bytes);
        if (rc0) return rc0;
    }
    SqRef *d_refs = static_cast<SqRef *>(dev);
    unsigned long long *d_io = reinterpret_cast<unsigned long long *>(static_cast<uint8_t *>(dev) +
                                                                      up256(sizeof(SqRef) * (N + 1)));
    int32_t *d_used = reinterpret_cast<int32_t *>(reinterpret_cast<uint8_t *>(d_io) + 64);
    int32_t *d_nused = d_used + S.K + 1;
    std::vector<unsigned long long> io = {slo, shi, (unsigned long long)v, 1ull};
    int rc = TDP_OK;
    int nused = 0;
    ok = true;
    while (c != end && ok) {
        // the stretch of steps back from cut c to the kept table at or before it (toward the sweep's start)
        int stop = c - dir;
        while (stop != end && !kept.at[stop].mem) stop -= dir;
        refs.clear();
        for (int x = c - dir;; x -= dir) {               // the table before each step, walking back
            Tab t;
            rc = table_at(R, dir, caps, use_h, kept, cache, x, t);
            if (rc) break;
            refs.push_back(SqRef{t.ents, t.slots, t.cap - 1, 0});
            if (x == stop) break;
        }
        if (rc) break;
        const int first = dir > 0 ? c - 1 : c;           // position of the first step undone
        if ((rc = hipMemcpyAsync(d_refs, refs.data(), sizeof(SqRef) * refs.size(), hipMemcpyHostToDevice, R.st)) ||
            (rc = hipMemcpyAsync(d_io, io.data(), 32, hipMemcpyHostToDevice, R.st)) ||
            (rc = hipMemcpyAsync(d_nused, &nused, 4, hipMemcpyHostToDevice, R.st))) {
            set_error("sparse query traceback: copy failed");
            rc = TDP_EHIP;
            break;
        }
        k_sq_walk<<<1, 64, 0, R.st>>>(d_refs, R.d_prm, first, (int)refs.size(), -dir, bits_word, d_io, d_used,
                                      d_nused, S.K + 1);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(io.data(), d_io, 32, hipMemcpyDeviceToHost, R.st) != hipSuccess ||
            hipMemcpyAsync(&nused, d_nused, 4, hipMemcpyDeviceToHost, R.st) != hipSuccess ||
            hipStreamSynchronize(R.st) != hipSuccess) {
            set_error("sparse query traceback failed on the device");
            rc = TDP_EHIP;
            break;
        }
        ok = io[3] != 0 && nused <= S.K + 1;
        c = stop;
    }
    if (rc == TDP_OK && ok) {
        std::vector<int32_t> got(std::max(nused, 1));
        if (hipMemcpyAsync(got.data(), d_used, 4ull * nused, hipMemcpyDeviceToHost, R.st) != hipSuccess ||
            hipStreamSynchronize(R.st) != hipSuccess) {
            set_error("sparse query traceback: read-back failed");
            rc = TDP_EHIP;
        } else {
            used.insert(used.end(), got.begin(), got.begin() + nused);
            ok = io[0] == 0 && io[1] == 0 && io[2] == 0;
        }
    }
    R.dfree(dev);
    return rc;
}

// The forward frontier's tables are kept like a plain sweep's: the traceback of the forward half recomputes a
// stretch as a plain sweep with caps a(c) - 1 would (the frontier keeps weights <= a(p) - 1 after step p, and where
// a drops the table is cut to a(c) - 1, which is the same table: a plain step of cap a(p) - 1 over the truncated
// table keeps exactly the states of value <= min(a(p), ...) - 1 that the frontier keeps).
static void frontier_caps(const tdp_sq_spec_t &S, std::vector<int32_t> &fc) {
    fc.assign(S.N + 1, 0);
    for (int c = 0; c <= S.N; c++) fc[c] = S.split[c] - 1;
}

}  // namespace tdp

using namespace tdp;

extern "C" {

int tdp_sq_run(const tdp_sq_spec_t *spec, uint64_t *fw_counts, uint64_t *bw_counts, uint64_t *ch_counts,
               int32_t *witness, tdp_sq_result_t *res, void *stream) {
    TDP_REQUIRE(spec && res && spec->N >= 0 && spec->prm && spec->tm && spec->every >= 1, "tdp_sq_run: bad spec");
    const tdp_sq_spec_t &S = *spec;
    const int N = S.N;
    const bool profile = S.mode == TDP_SQ_PROFILE_FW || S.mode == TDP_SQ_PROFILE_BW;
    // one sweep keeps weights <= K in 8-bit values; the meet in the middle splits K into two such parts
    TDP_REQUIRE(S.K >= 1 && S.K <= (S.mode == TDP_SQ_MITM ? 2 * 254 : 254), "tdp_sq_run: threshold %d", S.K);
    TDP_REQUIRE(S.mode == TDP_SQ_FORWARD || S.mode == TDP_SQ_MITM || profile, "tdp_sq_run: mode %d", S.mode);
    if (S.mode == TDP_SQ_MITM) {
        TDP_REQUIRE(S.split != nullptr, "tdp_sq_run: meet in the middle without a split");
        for (int c = 0; c <= N; c++) {
            TDP_REQUIRE(S.split[c] >= 1 && S.split[c] <= S.K && S.split[c] <= 254 && S.K - S.split[c] <= 254 &&
                            (c == 0 || S.split[c] <= S.split[c - 1]),
                        "tdp_sq_run: split must be non-increasing, within 1..K, with both parts <= 254 (cut %d)", c);
        }
    }
    TDP_REQUIRE(!S.use_h || (S.bnd_ncls[0] && S.bnd_ncls[1] && S.bnd_off[0] && S.bnd_off[1] && S.bnd_mask[0] &&
                             S.bnd_mask[1] && S.bnd_w[0] && S.bnd_w[1] && S.bnd_extra[0] && S.bnd_extra[1]),
                "tdp_sq_run: lookahead without bound data");
    memset(res, 0, sizeof(*res));
    res->value = -1;
    res->join_cut = -1;
    hipStream_t st = as_stream(stream);
    SqRun R(S, st);
    TDP_REQUIRE(!S.dev_alloc == !S.dev_free, "tdp_sq_run: dev_alloc and dev_free come together");
    {
        int rc0 = R.dalloc((void **)&R.d_cnt, 8 * 272);
        if (rc0) return rc0;
    }
    if (hipHostMalloc((void **)&R.h_cnt, 8 * 272, hipHostMallocDefault) != hipSuccess) {
        R.dfree(R.d_cnt);
        set_error("tdp_sq_run: pinned host counters");
        return TDP_EHIP;
    }
    // device copies of what kernels read: the column words (traceback) and the lookahead classes (steps)
    void *dconst = nullptr;
    uint64_t nclass[2] = {0, 0};
    if (S.use_h)
        for (int d = 0; d < 2; d++)
            for (int c = 0; c <= N; c++)
                nclass[d] = std::max<uint64_t>(nclass[d], (uint64_t)(S.bnd_off[d][c] + S.bnd_ncls[d][c]));
    const uint64_t prm_bytes = up256(80ull * std::max(N, 1));
    const uint64_t mb[2] = {up256(16 * std::max<uint64_t>(nclass[0], 1)), up256(16 * std::max<uint64_t>(nclass[1], 1))};
    const uint64_t wb[2] = {up256(8 * std::max<uint64_t>(nclass[0], 1)), up256(8 * std::max<uint64_t>(nclass[1], 1))};
    if (R.dalloc(&dconst, prm_bytes + mb[0] + mb[1] + wb[0] + wb[1])) {
        (void)hipStreamSynchronize(st);
        R.dfree(R.d_cnt);
        (void)hipHostFree(R.h_cnt);
        return TDP_ENOMEM;
    }
    {
        uint8_t *q = static_cast<uint8_t *>(dconst);
        R.d_prm = reinterpret_cast<uint64_t *>(q);
        q += prm_bytes;
        for (int d = 0; d < 2; d++) {
            R.d_bmask[d] = reinterpret_cast<uint64_t *>(q);
            q += mb[d];
            R.d_bw[d] = reinterpret_cast<int64_t *>(q);
            q += wb[d];
        }
        if (N) TDP_HIP_CHECK(hipMemcpyAsync(R.d_prm, S.prm, 80ull * N, hipMemcpyHostToDevice, st));
        for (int d = 0; d < 2 && S.use_h; d++)
            if (nclass[d]) {
                TDP_HIP_CHECK(hipMemcpyAsync(R.d_bmask[d], S.bnd_mask[d], 16 * nclass[d], hipMemcpyHostToDevice, st));
                TDP_HIP_CHECK(hipMemcpyAsync(R.d_bw[d], S.bnd_w[d], 8 * nclass[d], hipMemcpyHostToDevice, st));
            }
    }
    Kept bw, fw;
    std::vector<Tab> bcache, fcache, frontier_tabs;
    std::vector<int32_t> used;
    int rc = TDP_OK;
    void *chmem = nullptr;                     // the forward frontier's changed-state buffer
    auto finish = [&](int code) {
        R.dfree(chmem);
        chmem = nullptr;
        for (Tab &t : bw.at) R.drop(t);
        for (Tab &t : fw.at) R.drop(t);
        for (Tab &t : bcache) R.drop(t);
        for (Tab &t : fcache) R.drop(t);
        R.dfree(R.d_cnt);
        R.dfree(dconst);
        (void)hipStreamSynchronize(st);
        (void)hipHostFree(R.h_cnt);
        res->status = R.status;
        res->fail_cut = R.fail_cut;
        res->steps = R.steps;
        res->peak_states = R.peak_states;
        res->peak_bytes = R.peak_held;
        res->seconds = std::chrono::duration<double>(Clock::now() - R.t0).count();
        return code;
    };
    // ---- profiles: one plain sweep with histograms
    if (profile) {
        const int dir = S.mode == TDP_SQ_PROFILE_FW ? +1 : -1;
        std::vector<int32_t> caps(N + 1, S.K);
        TDP_REQUIRE(S.hist && S.hist_top >= 0 && S.hist_top <= 254, "tdp_sq_run: profile without a histogram");
        rc = plain_sweep(R, dir, caps.data(), false, dir > 0 ? fw : bw, dir > 0 ? fw_counts : bw_counts, S.hist,
                         S.hist_top);
        return finish(rc);
    }
    // ---- forward mode: one sweep keeping weights <= K, value at state 0 of cut N
    if (S.mode == TDP_SQ_FORWARD) {
        std::vector<int32_t> caps(N + 1, S.K);
        rc = plain_sweep(R, +1, caps.data(), S.use_h != 0, fw, fw_counts, nullptr, 0);
        if (rc || R.status != TDP_SQ_DONE) return finish(rc);
        res->forward_seconds = R.elapsed();
        // the value of state 0 at cut N: the last table has one state at most (no rows are active at cut N)
        const Tab &last = fw.at[N];
        TDP_HIP_CHECK(hipMemcpyAsync(R.h_cnt, last.ents, 16, hipMemcpyDeviceToHost, st));
        TDP_HIP_CHECK(hipStreamSynchronize(st));
        const int v = last.n ? (int)(R.h_cnt[1] >> 56) : 255;
        if (last.n && (R.h_cnt[0] | (R.h_cnt[1] & kKeyMask))) {
            set_error("sparse query: a nonzero state at the last cut");
            return finish(TDP_EINVAL);
        }
        if (v <= S.K) {
            res->value = v;
            res->join_cut = N;
            if (S.witness) {
                const auto t1 = Clock::now();
                bool ok = false;
                rc = walk_back(R, +1, caps.data(), S.use_h != 0, fw, fcache, N, 0, 0, v, used, ok);
                if (rc) return finish(rc);
                if (!ok || (int)used.size() != v) {
                    set_error("sparse query traceback lost the path");
                    return finish(TDP_EHIP);
                }
                res->traceback_seconds = std::chrono::duration<double>(Clock::now() - t1).count();
            }
        }
        res->nwitness = (int)used.size();
        for (size_t i = 0; i < used.size(); i++) witness[i] = used[i];
        return finish(TDP_OK);
    }
    // ---- meet in the middle: backward sweep with caps K - a(c)
    std::vector<int32_t> bcaps(N + 1), fcaps;
    for (int c = 0; c <= N; c++) bcaps[c] = S.K - S.split[c];
    frontier_caps(S, fcaps);
    rc = plain_sweep(R, -1, bcaps.data(), S.use_h != 0, bw, bw_counts, nullptr, 0);
    if (rc || R.status != TDP_SQ_DONE) return finish(rc);
    res->backward_seconds = R.elapsed();
    // forward frontier with joins
    fw.at.assign(N + 1, Tab());
    Tab cur;
    rc = R.start_table(cur);
    if (rc) return finish(rc);
    fw.at[0] = cur;
    fw.bytes = cur.bytes;
    bool cur_kept = true;
    if (fw_counts) fw_counts[0] = 1;
    if (ch_counts) ch_counts[0] = 1;
    int best = -1, best_cut = -1, best_fl = 0;
    uint64_t best_lo = 0, best_hi = 0, joined = 0;
    Ent *ch = nullptr;
    uint64_t chcap = 0;
    auto join = [&](const Ent *f, uint64_t nf, int c) -> int {
        if (!nf) return TDP_OK;
        Tab b;
        int r2 = table_at(R, -1, bcaps.data(), S.use_h != 0, bw, bcache, c, b);
        if (r2) return r2;
        if (!b.n) return TDP_OK;
        const uint64_t tmw[2] = {S.tm[2 * c], S.tm[2 * c + 1]};
        r2 = tdp_sp_join(f, nf, b.ents, b.slots, b.cap, tmw, R.d_cnt + 4, st);
        if (r2) return r2;
        TDP_HIP_CHECK(hipMemcpyAsync(R.h_cnt + 4, R.d_cnt + 4, 32, hipMemcpyDeviceToHost, st));
        TDP_HIP_CHECK(hipStreamSynchronize(st));
        joined += nf;
        const unsigned long long tot = R.h_cnt[4];
        if (tot != ~0ull && (best < 0 || (int)tot < best)) {
            best = (int)tot;
            best_cut = c;
            best_hi = R.h_cnt[5];
            best_lo = R.h_cnt[6];
            best_fl = (int)R.h_cnt[7];
        }
        return TDP_OK;
    };
    const int stop_at = std::max(S.lower, 0);
    rc = join(cur.ents, 1, 0);
    if (rc) return finish(rc);
    for (int p = 0; p < N && !(best >= 0 && best <= stop_at); p++) {
        const int c = p + 1;
        const uint64_t need = 2 * std::max<uint64_t>(cur.n, 1);
        if (chcap < need) {
            R.dfree(chmem);
            chmem = nullptr;
            chcap = std::max<uint64_t>(need, 2 * chcap);
            rc = R.dalloc(&chmem, 16 * chcap);
            if (rc) {
                set_error("sparse query: changed-state buffer of %llu entries", (unsigned long long)chcap);
                return finish(TDP_EHIP);
            }
            ch = static_cast<Ent *>(chmem);
        }
        Tab nxt;
        const int a = S.split[p];
        rc = R.step(cur, p, +1, a, a, S.use_h != 0, nxt, ch);
        if (rc) return finish(rc);
        rc = R.read_step();
        if (rc) return finish(rc);
        if (R.h_cnt[2]) {
            set_error("sparse query: step output past its capacity");
            return finish(TDP_EHIP);
        }
        nxt.n = R.h_cnt[0];
        const uint64_t nch = R.h_cnt[1];
        Tab extra;
        if (S.split[c] < a) {
            // the bound drops: join the whole table, then keep its states of value < a(c)
            extra = nxt;
            Tab cut;
            rc = R.alloc(cut, nxt.n, index_cap(std::max<uint64_t>(nxt.n, 1)));
            if (rc) return finish(rc);
            rc = tdp_sp_truncate(nxt.ents, nxt.n, S.split[c], cut.ents, R.d_cnt, st);
            if (rc) return finish(rc);
            rc = R.read(1);
            if (rc) return finish(rc);
            cut.n = R.h_cnt[0];
            k_sq_clear<<<grid_for(cut.cap, 4096), kSpBlock, 0, st>>>(cut.slots, cut.cap);
            if (cut.n) k_sq_index<<<grid_for(cut.n, ~0ull), kSpBlock, 0, st>>>(cut.ents, cut.n, cut.slots, cut.cap - 1);
            TDP_LAUNCH_CHECK();
            nxt = cut;
        }
        if (fw_counts) fw_counts[c] = nxt.n;
        if (ch_counts) ch_counts[c] = nch;
        rc = join(ch, nch, c);
        if (!rc && extra.mem) rc = join(extra.ents, extra.n, c);
        if (extra.mem) R.drop(extra);
        if (rc) return finish(rc);
        if (!cur_kept) R.drop(cur);
        cur = nxt;
        cur_kept = is_checkpoint(c, N, S.every) || fw.bytes + nxt.bytes <= S.keep_bytes;
        if (cur_kept) {
            fw.at[c] = nxt;
            fw.bytes += nxt.bytes;
        }
        if (R.check_budget(nxt.n, nch, c)) break;
    }
    if (!cur_kept) R.drop(cur);
    R.dfree(chmem);
    chmem = nullptr;
    res->joined = joined;
    res->forward_seconds = R.elapsed() - res->backward_seconds;
    if (R.status != TDP_SQ_DONE) return finish(TDP_OK);
    if (best >= 0 && best <= S.K) {
        res->value = best;
        res->join_cut = best_cut;
        if (S.witness) {
            const auto t1 = Clock::now();
            bool ok1 = false, ok2 = false;
            // forward half: a plain recomputation with caps a(c) - 1 rebuilds the frontier's tables
            rc = walk_back(R, +1, fcaps.data(), S.use_h != 0, fw, fcache, best_cut, best_lo, best_hi & kKeyMask,
                           best_fl, used, ok1);
            if (!rc) {
                const uint64_t rlo = best_lo ^ S.tm[2 * best_cut], rhi = (best_hi ^ S.tm[2 * best_cut + 1]) & kKeyMask;
                rc = walk_back(R, -1, bcaps.data(), S.use_h != 0, bw, bcache, best_cut, rlo, rhi, best - best_fl,
                               used, ok2);
            }
            if (rc) return finish(rc);
            if (!ok1 || !ok2 || (int)used.size() != best) {
                set_error("sparse query traceback lost the path");
                return finish(TDP_EHIP);
            }
            res->traceback_seconds = std::chrono::duration<double>(Clock::now() - t1).count();
        }
    }
    res->nwitness = (int)used.size();
    for (size_t i = 0; i < used.size(); i++) witness[i] = used[i];
    return finish(TDP_OK);
}

}  // extern "C"
