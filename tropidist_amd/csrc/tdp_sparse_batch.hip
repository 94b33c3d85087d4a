// Many meet-in-the-middle sparse queries at once, resident on one GPU (the query of cdist/sparse_dp.py:985-1095 with
// the default split, and the column step / join of cdist/sparse_dp.c:116-236, for every active query per launch).
//
// A job is one query: a trellis of N columns, threshold K and split a (forward frontier keeps weights <= a - 1 and
// joins its new or improved states of weight <= a; the backward sweep keeps weights <= K - a).  Its life:
//   backward start   table {0: 0} at cut N
//   backward steps   p = N-1 .. 0: table at cut p from the table at cut p + 1
//   forward start    table {0: 0} at cut 0, joined with the backward table at cut 0
//   forward steps    p = 0 .. N-1: table and changed states at cut p + 1; the changed states joined with the
//                    backward table at cut p + 1
// Every table is kept, with its index, until the job is released: the joins read the backward tables and the
// traceback walks both directions without recomputing a segment.
//
// One global step of the engine advances every active job by one column: one launch steps all jobs' tables, one
// read-back brings every job's counts (and the previous joins' least totals), then one launch clears the new tables'
// indexes and one inserts them and runs the joins.  Jobs enter and leave between global steps (continuous batching).
// The result of a job does not depend on the others: each owns its tables, counters and memory, and the kernels
// give every job the same blocks-to-entries mapping whatever else is in the launch.
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "tdp_sparse.h"

namespace tdp {

// ---------------------------------------------------------------------------------------------- launch records
struct SpbStep {                      // one job's column step
    const Ent *old;
    uint64_t n;
    const uint32_t *slots;
    uint64_t capmask;
    Ent *out;                         // capacity 2n entries
    Ent *ch;                          // changed states (capacity 2n), or null
    uint64_t upd_lo, upd_hi, sm_lo, sm_hi, sb_lo, sb_hi;
    int32_t cap, kb;                  // candidates from values < cap; the table keeps values < kb
    unsigned long long *cnt;          // [0] table count | changed count << 32, [2] overflow flag
    uint64_t pad0, pad1;
};
static_assert(sizeof(SpbStep) == 128, "SpbStep layout");

struct SpbPost {                      // one job's new table (index) and join
    Ent *ents;
    uint64_t n;
    uint32_t *slots;
    uint64_t capmask;
    const Ent *jf;                    // states to join (jn == 0: no join)
    uint64_t jn;
    const Ent *jb;                    // the backward table they meet, with its index
    const uint32_t *jslots;
    uint64_t jcapmask;
    uint64_t tm_lo, tm_hi;            // key(b) = key(f) ^ tmask
    unsigned long long *cnt;          // cleared for the job's next step; cnt[3] = least join total
    uint32_t start;                   // the new table is the start table {0: 0}
    uint32_t pad0;
    uint64_t pad1, pad2, pad3;
};
static_assert(sizeof(SpbPost) == 128, "SpbPost layout");

struct SpbTabRef {
    const Ent *ents;
    const uint32_t *slots;
    uint64_t capmask;
    uint64_t pad;
};

struct SpbTrace {
    const SpbTabRef *fw, *bw;         // per cut 0..N
    const uint64_t *prm;              // per column: 10 words (col, fwd mask, fwd bits, bwd mask, bwd bits; lo/hi)
    int32_t N, c;                     // the join cut
    uint64_t sl_lo, sl_hi;            // forward state and value at the cut
    uint64_t sr_lo, sr_hi;            // backward state and value at the cut
    int32_t fl, br;
    int32_t *pos;                     // used column positions (capacity fl + br)
    int32_t *result;                  // [0] count, or -1 when a lookup does not continue the path
    uint64_t pad;
};

constexpr int kPrmWords = 10;

// ---------------------------------------------------------------------------------------------- kernels
// Grids are njobs * per_job blocks; job j owns blocks [j * per_job, (j + 1) * per_job) and walks its entries with a
// stride of per_job * kSpBlock (per_job * kSpBlock * kSpTile for the step's tiles).  Loop bounds depend on the block
// only, so every thread of a block reaches the block-wide appends and reductions together.
__global__ __launch_bounds__(kSpBlock) void k_spb_step(const SpbStep *__restrict__ jobs, uint32_t per_job) {
    const SpbStep &J = jobs[blockIdx.x / per_job];
    StepArgs a;
    a.upd_lo = J.upd_lo;
    a.upd_hi = J.upd_hi;
    a.sm_lo = J.sm_lo;
    a.sm_hi = J.sm_hi;
    a.sb_lo = J.sb_lo;
    a.sb_hi = J.sb_hi;
    a.cap = J.cap;
    a.kb = J.kb;
    a.want_changed = J.ch != nullptr;
    a.out_cap = 2 * J.n;
    a.ch_cap = 2 * J.n;
    BoundArgs b;
    b.mask = nullptr;
    b.w = nullptr;
    b.ncls = 0;
    b.on = 0;
    b.extra = 0;
    b.tm_lo = b.tm_hi = 0;
    b.K = 0;
    const uint64_t tile = (uint64_t)kSpBlock * kSpTile;
    for (uint64_t base = (uint64_t)(blockIdx.x % per_job) * tile; base < J.n; base += (uint64_t)per_job * tile)
        sp_step_tile<false>(J.old, J.n, base, J.slots, J.capmask, a, b, J.out, J.ch, &J.cnt[0], &J.cnt[2], nullptr,
                            0);
}

__global__ __launch_bounds__(kSpBlock) void k_spb_fill(const SpbPost *__restrict__ jobs, uint32_t per_job) {
    const SpbPost &J = jobs[blockIdx.x / per_job];
    const uint32_t x = blockIdx.x % per_job;
    if (x == 0 && threadIdx.x == 0) {
        if (J.jn) J.cnt[3] = ~0ull;
        if (J.start) J.ents[0] = Ent{0, 0};
    }
    const uint64_t stride = (uint64_t)per_job * kSpBlock;
    for (uint64_t i = (uint64_t)x * kSpBlock + threadIdx.x; i <= J.capmask; i += stride) J.slots[i] = kEmpty;
}

__global__ __launch_bounds__(kSpBlock) void k_spb_post(const SpbPost *__restrict__ jobs, uint32_t per_job) {
    const SpbPost &J = jobs[blockIdx.x / per_job];
    const uint32_t x = blockIdx.x % per_job;
    if (x == 0 && threadIdx.x < 3) J.cnt[threadIdx.x] = 0;
    const uint64_t stride = (uint64_t)per_job * kSpBlock;
    for (uint64_t i = (uint64_t)x * kSpBlock + threadIdx.x; i < J.n; i += stride) sp_insert(J.ents, i, J.slots, J.capmask);
    // the joins: the block's least total, one atomic per block on the job's least
    uint32_t t = ~0u;
    for (uint64_t i = (uint64_t)x * kSpBlock + threadIdx.x; i < J.jn; i += stride) {
        Ent f = J.jf[i];
        uint32_t vb = sp_find(J.jb, J.jslots, J.jcapmask, f.lo ^ J.tm_lo, (f.hv & kKeyMask) ^ J.tm_hi);
        if (vb != 255u) t = min(t, (uint32_t)(f.hv >> 56) + vb);
    }
    if (J.jn) block_min_atomic(t, &J.cnt[3]);
}

// One step back along the path (cdist/sparse_dp.py:812-819): the state before the column is s | bits; it stayed if
// the table holds it with value v, else it came from s | bits ^ upd with value v - 1 (the column is used).
__device__ bool spb_undo(const SpbTabRef &t, uint64_t &slo, uint64_t &shi, int &v, uint64_t blo, uint64_t bhi,
                         uint64_t ulo, uint64_t uhi, bool &used) {
    uint64_t mlo = slo | blo, mhi = shi | bhi;
    if ((int)sp_find(t.ents, t.slots, t.capmask, mlo, mhi) == v) {
        slo = mlo;
        shi = mhi;
        used = false;
        return true;
    }
    if (!(ulo | uhi) || (int)sp_find(t.ents, t.slots, t.capmask, mlo ^ ulo, mhi ^ uhi) != v - 1) return false;
    slo = mlo ^ ulo;
    shi = mhi ^ uhi;
    v -= 1;
    used = true;
    return true;
}

// One thread per traced job: the forward half back from the join cut to cut 0, then the backward half forward to
// cut N.  Both must end in state 0 with value 0.
__global__ void k_spb_trace(const SpbTrace *__restrict__ jobs, int njobs) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= njobs) return;
    const SpbTrace &T = jobs[j];
    int count = 0;
    bool ok = true, used = false;
    uint64_t slo = T.sl_lo, shi = T.sl_hi;
    int v = T.fl;
    for (int p = T.c - 1; p >= 0 && ok; p--) {
        const uint64_t *q = T.prm + (uint64_t)p * kPrmWords;
        ok = spb_undo(T.fw[p], slo, shi, v, q[4], q[5], q[0], q[1], used);
        if (ok && used) {
            if (count < T.fl + T.br) T.pos[count] = p;
            count++;
        }
    }
    ok = ok && slo == 0 && shi == 0 && v == 0;
    slo = T.sr_lo;
    shi = T.sr_hi;
    v = T.br;
    for (int p = T.c; p < T.N && ok; p++) {
        const uint64_t *q = T.prm + (uint64_t)p * kPrmWords;
        ok = spb_undo(T.bw[p + 1], slo, shi, v, q[8], q[9], q[0], q[1], used);
        if (ok && used) {
            if (count < T.fl + T.br) T.pos[count] = p;
            count++;
        }
    }
    ok = ok && slo == 0 && shi == 0 && v == 0 && count == T.fl + T.br;
    T.result[0] = ok ? count : -1;
}

// ---------------------------------------------------------------------------------------------- the engine
namespace {

constexpr uint64_t kAlign = 256;
constexpr uint64_t kMaxGridBlocks = 1ull << 20;

uint64_t align_up(uint64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over one device arena (free blocks coalesced on release).
class Arena {
public:
    uint8_t *base = nullptr;
    uint64_t size = 0;
    std::map<uint64_t, uint64_t> free_;     // offset -> bytes

    void init(uint8_t *b, uint64_t s) {
        base = b;
        size = s;
        free_.clear();
        free_[0] = s;
    }
    uint64_t in_use = 0, peak = 0;

    uint8_t *alloc(uint64_t bytes) {
        bytes = align_up(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            uint64_t off = it->first, left = it->second - bytes;
            free_.erase(it);
            if (left) free_[off + bytes] = left;
            in_use += bytes;
            peak = std::max(peak, in_use);
            return base + off;
        }
        return nullptr;
    }
    void release(uint8_t *p, uint64_t bytes) {
        bytes = align_up(bytes ? bytes : 1);
        in_use -= bytes;
        uint64_t off = (uint64_t)(p - base);
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }
};

struct Tab {
    Ent *ents = nullptr;
    uint64_t n = 0;
    uint32_t *slots = nullptr;
    uint64_t capmask = 0;
};

enum Phase { kBwStart, kBw, kFwStart, kFw, kFwLast, kDone };

struct JobRec {
    int64_t id = 0;
    int N = 0, K = 0, a = 0;
    int phase = kBwStart, p = 0;
    int status = TDP_SPB_RUNNING;
    int slot = -1;
    uint64_t max_states = 0;
    std::vector<uint64_t> prm;        // N x 10
    std::vector<uint64_t> tm;         // (N + 1) x 2
    std::vector<Tab> fw, bw;          // per cut
    // memory: segments of the arena, bump-allocated
    std::vector<std::pair<uint8_t *, uint64_t>> segs;
    uint8_t *bump = nullptr, *bump_end = nullptr;
    uint64_t last_seg = 0;
    Ent *ch[2] = {nullptr, nullptr};
    uint64_t chcap[2] = {0, 0};
    // the step in flight
    bool stepped = false;
    Ent *out = nullptr;
    uint64_t out_bytes = 0;
    uint32_t *out_slots = nullptr;    // the new table's index, carved from the step's reservation
    int par = 0;                      // changed buffer of the step
    uint64_t nout = 0, nch = 0;
    // the join in flight (read at the next global step)
    bool joined = false;
    const Ent *jf = nullptr;
    uint64_t jn = 0;
    int jcut = 0;
    // the best join so far (strictly better totals replace it; the earliest cut wins a tie)
    int best = -1, best_cut = -1, best_fl = 0;
    uint64_t best_lo = 0, best_hi = 0;
    uint64_t peak = 1, state_steps = 0;
    uint64_t held = 0, peak_held = 0;  // arena bytes of the job's segments and changed buffers
};

}  // namespace

class SpBatch {
public:
    hipStream_t st = nullptr;
    int max_jobs = 0;
    Arena arena;
    uint8_t *d_arena = nullptr;
    SpbStep *d_step = nullptr, *h_step = nullptr;
    SpbPost *d_post = nullptr, *h_post = nullptr;
    unsigned long long *d_cnt = nullptr, *h_cnt = nullptr;
    unsigned long long *d_best = nullptr, *h_best = nullptr;
    std::vector<JobRec *> jobs;       // active and finished (until released), in order of admission
    std::vector<int> free_slots;
    std::vector<int64_t> finished;   // ids that finished (or failed) since the last poll
    int64_t next_id = 1;
    uint64_t global_steps = 0;
    std::mutex mu;                    // one caller at a time per handle

    ~SpBatch() {
        for (JobRec *j : jobs) delete j;
        if (d_arena) (void)hipFree(d_arena);
        if (d_step) (void)hipFree(d_step);
        if (d_post) (void)hipFree(d_post);
        if (d_cnt) (void)hipFree(d_cnt);
        if (d_best) (void)hipFree(d_best);
        if (h_step) (void)hipHostFree(h_step);
        if (h_post) (void)hipHostFree(h_post);
        if (h_cnt) (void)hipHostFree(h_cnt);
        if (h_best) (void)hipHostFree(h_best);
    }

    int init(uint64_t arena_bytes, int maxj, hipStream_t s) {
        st = s;
        max_jobs = maxj;
        TDP_HIP_CHECK(hipMalloc((void **)&d_arena, arena_bytes));
        arena.init(d_arena, arena_bytes);
        TDP_HIP_CHECK(hipMalloc((void **)&d_step, sizeof(SpbStep) * maxj));
        TDP_HIP_CHECK(hipMalloc((void **)&d_post, sizeof(SpbPost) * maxj));
        TDP_HIP_CHECK(hipMalloc((void **)&d_cnt, 32ull * maxj));
        TDP_HIP_CHECK(hipMalloc((void **)&d_best, 64));
        TDP_HIP_CHECK(hipHostMalloc((void **)&h_step, sizeof(SpbStep) * maxj, hipHostMallocDefault));
        TDP_HIP_CHECK(hipHostMalloc((void **)&h_post, sizeof(SpbPost) * maxj, hipHostMallocDefault));
        TDP_HIP_CHECK(hipHostMalloc((void **)&h_cnt, 32ull * maxj, hipHostMallocDefault));
        TDP_HIP_CHECK(hipHostMalloc((void **)&h_best, 64, hipHostMallocDefault));
        for (int s2 = maxj - 1; s2 >= 0; s2--) free_slots.push_back(s2);
        return TDP_OK;
    }

    JobRec *find(int64_t id) {
        for (JobRec *j : jobs)
            if (j->id == id) return j;
        return nullptr;
    }

    uint8_t *jalloc(JobRec *j, uint64_t bytes) {
        bytes = align_up(bytes ? bytes : 1);
        if (j->bump && (uint64_t)(j->bump_end - j->bump) >= bytes) {
            uint8_t *p = j->bump;
            j->bump += bytes;
            return p;
        }
        // segments double from 64 KiB (most queries are small) up to 256 MiB
        uint64_t want = std::max<uint64_t>(bytes, std::min<uint64_t>(std::max<uint64_t>(2 * j->last_seg, 64ull << 10),
                                                                     256ull << 20));
        uint8_t *seg = arena.alloc(want);
        if (!seg && want > bytes) seg = arena.alloc(want = bytes);
        if (!seg) return nullptr;
        j->held += align_up(want);
        j->peak_held = std::max(j->peak_held, j->held);
        j->segs.push_back({seg, want});
        j->last_seg = want;
        j->bump = seg + bytes;
        j->bump_end = seg + want;
        return seg;
    }

    void release_memory(JobRec *j) {
        for (auto &s : j->segs) arena.release(s.first, s.second);
        j->segs.clear();
        j->bump = j->bump_end = nullptr;
        for (int k = 0; k < 2; k++) {
            if (j->ch[k]) arena.release((uint8_t *)j->ch[k], 16 * j->chcap[k]);
            j->ch[k] = nullptr;
            j->chcap[k] = 0;
        }
        j->held = 0;
        j->fw.clear();
        j->bw.clear();
    }

    uint64_t nomem = 0;
    uint64_t nomem_last = 0;          // jobs that found the arena full at the last step

    void fail(JobRec *j, int status) {
        if (status == TDP_SPB_NOMEM) nomem++;
        j->status = status;
        j->phase = kDone;
        finished.push_back(j->id);
        j->stepped = j->joined = false;
        release_memory(j);
    }

    static uint64_t index_cap(uint64_t n) {
        uint64_t cap = 64;
        while (cap < 2 * n) cap <<= 1;
        return cap;
    }

    // Picks the least forward key among the changed states reaching the least total at the job's last join (the
    // tie rule of cdist/sparse_dp.c:207-236), synchronously.
    int pick(JobRec *j, int total) {
        const Tab &b = j->bw[j->jcut];
        uint64_t tmask[2] = {j->tm[2 * j->jcut], j->tm[2 * j->jcut + 1]};
        int rc = tdp_sp_join(j->jf, j->jn, b.ents, b.slots, b.capmask + 1, tmask, d_best, st);
        if (rc) return rc;
        TDP_HIP_CHECK(hipMemcpyAsync(h_best, d_best, 32, hipMemcpyDeviceToHost, st));
        TDP_HIP_CHECK(hipStreamSynchronize(st));
        if ((int)h_best[0] != total) {
            set_error("join pick disagrees with the join total (%d vs %d)", (int)h_best[0], total);
            return TDP_EHIP;
        }
        j->best = total;
        j->best_cut = j->jcut;
        j->best_hi = h_best[1];
        j->best_lo = h_best[2];
        j->best_fl = (int)h_best[3];
        return TDP_OK;
    }

    int add(const uint64_t *prm, const uint64_t *tm, int N, int K, int a, uint64_t max_states, int64_t *id) {
        TDP_REQUIRE(N >= 0 && K >= 1 && K <= 254 && a >= 1 && a <= K, "job N=%d K=%d split %d", N, K, a);
        TDP_REQUIRE(!free_slots.empty(), "all %d job slots are in use", max_jobs);
        JobRec *j = new JobRec();
        j->id = next_id++;
        j->N = N;
        j->K = K;
        j->a = a;
        j->max_states = max_states ? max_states : ~0ull;
        j->prm.assign(prm, prm + (size_t)N * kPrmWords);
        j->tm.assign(tm, tm + (size_t)(N + 1) * 2);
        for (size_t k = 1; k < j->prm.size(); k += 2) j->prm[k] &= kKeyMask;   // keys are 120 bits
        for (size_t k = 1; k < j->tm.size(); k += 2) j->tm[k] &= kKeyMask;
        j->fw.resize(N + 1);
        j->bw.resize(N + 1);
        j->slot = free_slots.back();
        free_slots.pop_back();
        j->phase = kBwStart;
        jobs.push_back(j);
        *id = j->id;
        return TDP_OK;
    }

    int step(int *active) {
        const uint64_t nomem_before = nomem;
        // ---- 1. column steps of every stepping job
        uint32_t nstep = 0;
        uint64_t maxn = 1;
        for (JobRec *j : jobs) {
            j->stepped = false;
            if (j->phase != kBw && j->phase != kFw) continue;
            const bool fwd = j->phase == kFw;
            const Tab &cur = fwd ? j->fw[j->p] : j->bw[j->p + 1];
            const uint64_t n = cur.n;
            const uint64_t n1 = std::max<uint64_t>(n, 1);
            if (n >= (1ull << 31)) { fail(j, TDP_SPB_STATES); continue; }       // the step counts in 32 bits
            Ent *ch = nullptr;
            const int par = (int)(global_steps & 1);
            if (fwd) {
                if (j->chcap[par] < 2 * n1) {
                    if (j->ch[par]) {
                        arena.release((uint8_t *)j->ch[par], 16 * j->chcap[par]);
                        j->held -= align_up(16 * j->chcap[par]);
                    }
                    uint64_t cap = 2 * std::max<uint64_t>(n, 64);
                    j->ch[par] = (Ent *)arena.alloc(16 * cap);
                    j->chcap[par] = j->ch[par] ? cap : 0;
                    if (!j->ch[par]) { fail(j, TDP_SPB_NOMEM); continue; }
                    j->held += align_up(16 * cap);
                    j->peak_held = std::max(j->peak_held, j->held);
                }
                ch = j->ch[par];
            }
            // one reservation for the step's output (up to 2n entries) and the new table's index
            const uint64_t out_part = align_up(32 * n1);
            uint8_t *out = jalloc(j, out_part + 4 * index_cap(2 * n1));
            if (!out) { fail(j, TDP_SPB_NOMEM); continue; }
            const uint64_t *q = j->prm.data() + (size_t)j->p * kPrmWords;
            SpbStep &S = h_step[nstep++];
            S.old = cur.ents;
            S.n = n;
            S.slots = cur.slots;
            S.capmask = cur.capmask;
            S.out = (Ent *)out;
            S.ch = ch;
            S.upd_lo = q[0];
            S.upd_hi = q[1];
            S.sm_lo = fwd ? q[2] : q[6];
            S.sm_hi = fwd ? q[3] : q[7];
            S.sb_lo = fwd ? q[4] : q[8];
            S.sb_hi = fwd ? q[5] : q[9];
            S.cap = fwd ? j->a : j->K - j->a;
            S.kb = fwd ? j->a : j->K - j->a + 1;
            S.cnt = d_cnt + 4 * (uint64_t)j->slot;
            S.pad0 = S.pad1 = 0;
            j->stepped = true;
            j->out = (Ent *)out;
            j->out_bytes = align_up(out_part + 4 * index_cap(2 * n1));
            j->par = par;
            maxn = std::max(maxn, n);
        }
        int max_slot = -1;
        for (JobRec *j : jobs)
            if (j->phase != kDone) max_slot = std::max(max_slot, j->slot);
        if (nstep) {
            TDP_HIP_CHECK(hipMemcpyAsync(d_step, h_step, sizeof(SpbStep) * nstep, hipMemcpyHostToDevice, st));
            const uint64_t tile = (uint64_t)kSpBlock * kSpTile;
            uint64_t per = std::min<uint64_t>((maxn + tile - 1) / tile, std::max<uint64_t>(1, kMaxGridBlocks / nstep));
            k_spb_step<<<(unsigned)(nstep * per), kSpBlock, 0, st>>>(d_step, (uint32_t)per);
            TDP_LAUNCH_CHECK();
        }
        if (max_slot >= 0)
            TDP_HIP_CHECK(hipMemcpyAsync(h_cnt, d_cnt, 32ull * (max_slot + 1), hipMemcpyDeviceToHost, st));
        TDP_HIP_CHECK(hipStreamSynchronize(st));
        global_steps++;

        // ---- 2. counts, budgets and the joins launched at the previous global step
        for (JobRec *j : jobs) {
            if (j->phase == kDone) continue;
            const unsigned long long *c = h_cnt + 4 * (uint64_t)j->slot;
            if (j->joined) {
                j->joined = false;
                unsigned long long tot = c[3];
                if (tot != ~0ull && (j->best < 0 || (int)tot < j->best) && pick(j, (int)tot) != TDP_OK) {
                    // this job alone fails (its error stays in tdp_last_error); the step goes on for the others, so
                    // every stepped job still gets its new table indexed and its counters cleared below
                    fail(j, TDP_SPB_FAILED);
                    continue;
                }
            }
            if (j->phase == kFwLast) {
                j->phase = kDone;
                j->status = TDP_SPB_DONE;
                finished.push_back(j->id);
                continue;
            }
            if (!j->stepped) continue;
            if (c[2]) {
                set_error("sparse batch: step output past its capacity (job %lld)", (long long)j->id);
                fail(j, TDP_SPB_FAILED);
                continue;
            }
            j->nout = c[0] & 0xFFFFFFFFull;
            j->nch = c[0] >> 32;
            j->state_steps += j->nout;
            j->peak = std::max<uint64_t>(j->peak, std::max<uint64_t>(j->nout, j->nch));
            if (std::max<uint64_t>(j->nout, j->nch) > j->max_states) { fail(j, TDP_SPB_STATES); continue; }
            // the table's index follows its entries inside the reservation (16 nout <= 32 n and
            // index_cap(nout) <= index_cap(2 n)); the unused tail is given back when the reservation is the job's
            // last bump allocation
            uint8_t *slots = (uint8_t *)j->out + align_up(16 * std::max<uint64_t>(j->nout, 1));
            j->out_slots = (uint32_t *)slots;
            uint8_t *end = slots + align_up(4 * index_cap(j->nout));
            if ((uint8_t *)j->out + j->out_bytes == j->bump) j->bump = end;
        }

        // ---- 3. new tables: index, and the forward joins
        uint32_t npost = 0;
        uint64_t maxcap = 64, maxw = 1;
        for (JobRec *j : jobs) {
            const bool start = j->phase == kBwStart || j->phase == kFwStart;
            if (!start && !j->stepped) continue;
            if (j->phase == kDone) continue;
            Tab t;
            int cut;
            uint64_t cap = 64;
            if (start) {
                uint8_t *m = jalloc(j, align_up(16) + 4 * index_cap(1));
                if (!m) { fail(j, TDP_SPB_NOMEM); continue; }
                t.ents = (Ent *)m;
                t.slots = (uint32_t *)(m + align_up(16));
                t.n = 1;
                cap = index_cap(1);
                cut = j->phase == kBwStart ? j->N : 0;
            } else {
                t.ents = j->out;
                t.n = j->nout;
                cap = index_cap(t.n);
                t.slots = j->out_slots;
                cut = j->phase == kBw ? j->p : j->p + 1;
            }
            t.capmask = cap - 1;
            SpbPost &P = h_post[npost++];
            memset(&P, 0, sizeof(P));
            P.ents = t.ents;
            P.n = t.n;
            P.slots = t.slots;
            P.capmask = t.capmask;
            P.cnt = d_cnt + 4 * (uint64_t)j->slot;
            P.start = start ? 1u : 0u;
            bool fwd = j->phase == kFwStart || j->phase == kFw;
            if (fwd) {
                j->fw[cut] = t;
                const Tab &b = j->bw[cut];
                const Ent *jf = start ? t.ents : j->ch[j->par];
                uint64_t jn = start ? 1 : j->nch;
                if (jn) {
                    P.jf = jf;
                    P.jn = jn;
                    P.jb = b.ents;
                    P.jslots = b.slots;
                    P.jcapmask = b.capmask;
                    P.tm_lo = j->tm[2 * cut];
                    P.tm_hi = j->tm[2 * cut + 1];
                    j->joined = true;
                    j->jf = jf;
                    j->jn = jn;
                    j->jcut = cut;
                }
            } else {
                j->bw[cut] = t;
            }
            maxcap = std::max(maxcap, cap);
            maxw = std::max(maxw, std::max<uint64_t>(t.n, P.jn));
            // advance
            switch (j->phase) {
                case kBwStart: j->phase = j->N ? kBw : kFwStart; j->p = j->N - 1; break;
                case kBw: if (j->p == 0) j->phase = kFwStart; else j->p--; break;
                case kFwStart: j->phase = j->N ? kFw : kFwLast; j->p = 0; break;
                case kFw: if (++j->p == j->N) j->phase = kFwLast; break;
                default: break;
            }
        }
        if (npost) {
            TDP_HIP_CHECK(hipMemcpyAsync(d_post, h_post, sizeof(SpbPost) * npost, hipMemcpyHostToDevice, st));
            uint64_t lim = std::max<uint64_t>(1, kMaxGridBlocks / npost);
            uint64_t perf = std::min<uint64_t>((maxcap + kSpBlock - 1) / kSpBlock, lim);
            k_spb_fill<<<(unsigned)(npost * perf), kSpBlock, 0, st>>>(d_post, (uint32_t)perf);
            TDP_LAUNCH_CHECK();
            uint64_t perp = std::min<uint64_t>((maxw + kSpBlock - 1) / kSpBlock, lim);
            k_spb_post<<<(unsigned)(npost * perp), kSpBlock, 0, st>>>(d_post, (uint32_t)perp);
            TDP_LAUNCH_CHECK();
        }
        nomem_last = nomem - nomem_before;
        int na = 0;
        for (JobRec *j : jobs) na += j->phase != kDone;
        *active = na;
        return TDP_OK;
    }

    int witness(JobRec *j, int32_t *pos, int cap, int *n) {
        TDP_REQUIRE(j->status == TDP_SPB_DONE && j->best >= 0 && j->best <= j->K,
                    "job %lld has no value to trace", (long long)j->id);
        const int fl = j->best_fl, br = j->best - j->best_fl, c = j->best_cut, N = j->N;
        TDP_REQUIRE(cap >= fl + br, "witness buffer of %d for %d columns", cap, fl + br);
        // device records: tables of both directions, column words, the trace record and its outputs
        uint64_t b_tabs = align_up(sizeof(SpbTabRef) * 2 * (N + 1)), b_prm = align_up(8ull * kPrmWords * std::max(N, 1));
        uint64_t total = b_tabs + b_prm + align_up(sizeof(SpbTrace)) + align_up(4ull * (fl + br + 2));
        std::vector<uint8_t> host(total, 0);
        SpbTabRef *tabs = (SpbTabRef *)host.data();
        for (int k = 0; k <= N; k++) {
            tabs[k] = SpbTabRef{j->fw[k].ents, j->fw[k].slots, j->fw[k].capmask, 0};
            tabs[N + 1 + k] = SpbTabRef{j->bw[k].ents, j->bw[k].slots, j->bw[k].capmask, 0};
        }
        if (N) memcpy(host.data() + b_tabs, j->prm.data(), 8ull * kPrmWords * N);
        uint8_t *dev = arena.alloc(total);
        const bool own = dev == nullptr;         // arena full: a device allocation of its own
        if (own) TDP_HIP_CHECK(hipMalloc((void **)&dev, total));
        SpbTrace T;
        memset(&T, 0, sizeof(T));
        T.fw = (const SpbTabRef *)dev;
        T.bw = (const SpbTabRef *)dev + (N + 1);
        T.prm = (const uint64_t *)(dev + b_tabs);
        T.N = N;
        T.c = c;
        T.sl_lo = j->best_lo;
        T.sl_hi = j->best_hi;
        T.fl = fl;
        T.sr_lo = j->best_lo ^ j->tm[2 * c];
        T.sr_hi = j->best_hi ^ j->tm[2 * c + 1];
        T.br = br;
        uint8_t *outp = dev + b_tabs + b_prm + align_up(sizeof(SpbTrace));
        T.result = (int32_t *)outp;
        T.pos = (int32_t *)outp + 1;
        memcpy(host.data() + b_tabs + b_prm, &T, sizeof(T));
        int rc = TDP_OK;
        std::vector<int32_t> res(fl + br + 1, 0);
        hipError_t e = hipMemcpyAsync(dev, host.data(), total, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            k_spb_trace<<<1, 64, 0, st>>>((const SpbTrace *)(dev + b_tabs + b_prm), 1);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(res.data(), outp, 4ull * (fl + br + 1), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (own) (void)hipFree(dev);
        else arena.release(dev, total);
        if (e != hipSuccess) {
            set_error("sparse batch traceback: %s", hipGetErrorString(e));
            return TDP_EHIP;
        }
        if (res[0] != fl + br) {
            set_error("sparse batch traceback lost the path (job %lld)", (long long)j->id);
            rc = TDP_EHIP;
        }
        *n = res[0] < 0 ? 0 : res[0];
        for (int k = 0; k < *n && k < cap; k++) pos[k] = res[1 + k];
        return rc;
    }

    void release(JobRec *j) {
        release_memory(j);
        free_slots.push_back(j->slot);
        jobs.erase(std::find(jobs.begin(), jobs.end(), j));
        delete j;
    }
};

}  // namespace tdp

using namespace tdp;

extern "C" {

void *tdp_spb_create(uint64_t arena_bytes, int max_jobs, void *stream) {
    if (arena_bytes < (1ull << 20) || max_jobs < 1 || max_jobs > (1 << 20)) {
        set_error("sparse batch: arena of %llu bytes, %d jobs", (unsigned long long)arena_bytes, max_jobs);
        return nullptr;
    }
    SpBatch *b = new SpBatch();
    if (b->init(arena_bytes, max_jobs, as_stream(stream)) != TDP_OK) {
        delete b;
        return nullptr;
    }
    return b;
}

int tdp_spb_free(void *h) {
    if (!h) return TDP_OK;
    SpBatch *b = static_cast<SpBatch *>(h);
    TDP_HIP_CHECK(hipStreamSynchronize(b->st));
    delete b;
    return TDP_OK;
}

int tdp_spb_add(void *h, const uint64_t *prm, const uint64_t *tmask, int32_t N, int32_t K, int32_t split,
                uint64_t max_states, int64_t *id) {
    TDP_REQUIRE(h && (prm || !N) && tmask && id, "tdp_spb_add: null argument");
    SpBatch *b = static_cast<SpBatch *>(h);
    std::lock_guard<std::mutex> g(b->mu);
    return b->add(prm, tmask, N, K, split, max_states, id);
}

int tdp_spb_step(void *h, int32_t *active) {
    TDP_REQUIRE(h && active, "tdp_spb_step: null argument");
    SpBatch *b = static_cast<SpBatch *>(h);
    std::lock_guard<std::mutex> g(b->mu);
    return b->step(active);
}

int tdp_spb_stats(void *h, uint64_t *out) {
    TDP_REQUIRE(h && out, "tdp_spb_stats: null argument");
    SpBatch *b = static_cast<SpBatch *>(h);
    std::lock_guard<std::mutex> g(b->mu);
    int64_t active = 0;
    for (JobRec *j : b->jobs) active += j->phase != kDone;
    out[0] = b->arena.size;
    out[1] = b->arena.in_use;
    out[2] = b->arena.peak;
    out[3] = b->nomem;
    out[4] = b->global_steps;
    out[5] = (uint64_t)active;
    out[6] = b->nomem_last;
    return TDP_OK;
}

int tdp_spb_poll(void *h, int64_t *ids, int32_t cap, int32_t *n) {
    TDP_REQUIRE(h && n && (ids || cap == 0) && cap >= 0, "tdp_spb_poll: bad arguments");
    SpBatch *b = static_cast<SpBatch *>(h);
    std::lock_guard<std::mutex> g(b->mu);
    int k = std::min<int>(cap, (int)b->finished.size());
    for (int i = 0; i < k; i++) ids[i] = b->finished[i];
    b->finished.erase(b->finished.begin(), b->finished.begin() + k);
    *n = k;
    return TDP_OK;
}

int tdp_spb_result(void *h, int64_t id, int32_t *status, int32_t *value, uint64_t *stats) {
    TDP_REQUIRE(h && status && value, "tdp_spb_result: null argument");
    SpBatch *b = static_cast<SpBatch *>(h);
    std::lock_guard<std::mutex> g(b->mu);
    JobRec *j = b->find(id);
    TDP_REQUIRE(j != nullptr, "no sparse batch job %lld", (long long)id);
    *status = j->status;
    *value = (j->status == TDP_SPB_DONE && j->best >= 0 && j->best <= j->K) ? j->best : -1;
    if (stats) {
        stats[0] = j->state_steps;
        stats[1] = j->peak;
        stats[2] = (uint64_t)(j->best_cut < 0 ? 0 : j->best_cut);
        stats[3] = j->peak_held;
    }
    return TDP_OK;
}

int tdp_spb_witness(void *h, int64_t id, int32_t *positions, int32_t cap, int32_t *n) {
    TDP_REQUIRE(h && positions && n, "tdp_spb_witness: null argument");
    SpBatch *b = static_cast<SpBatch *>(h);
    std::lock_guard<std::mutex> g(b->mu);
    JobRec *j = b->find(id);
    TDP_REQUIRE(j != nullptr, "no sparse batch job %lld", (long long)id);
    return b->witness(j, positions, cap, n);
}

int tdp_spb_cancel(void *h, int64_t id) {
    TDP_REQUIRE(h, "tdp_spb_cancel: null handle");
    SpBatch *b = static_cast<SpBatch *>(h);
    std::lock_guard<std::mutex> g(b->mu);
    JobRec *j = b->find(id);
    TDP_REQUIRE(j != nullptr, "no sparse batch job %lld", (long long)id);
    if (j->phase != kDone) b->fail(j, TDP_SPB_CANCELLED);
    return TDP_OK;
}

int tdp_spb_release(void *h, int64_t id) {
    TDP_REQUIRE(h, "tdp_spb_release: null handle");
    SpBatch *b = static_cast<SpBatch *>(h);
    std::lock_guard<std::mutex> g(b->mu);
    JobRec *j = b->find(id);
    TDP_REQUIRE(j != nullptr, "no sparse batch job %lld", (long long)id);
    TDP_REQUIRE(j->phase == kDone, "sparse batch job %lld is still running", (long long)id);
    b->release(j);
    return TDP_OK;
}

}  // extern "C"
