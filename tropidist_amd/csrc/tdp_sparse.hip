// Weight-truncated sparse trellis DP on gfx950 (the step / join / lookup semantics of cdist/sparse_dp.c:116-242 and
// cdist/sparse_dp.py:338-445, on device-resident tables).
//
// Table: n entries of 16 bytes, {lo, hv}: key = (hv & 2^56-1) << 64 | lo (up to 120 row slots), value = hv >> 56,
// plus an open-addressing index of uint32 entry numbers (power-of-two capacity >= 2n, linear probing, 0xFFFFFFFF =
// empty).  Keys in a table are distinct; the entry order is arbitrary (results never depend on it).
//
// One column step (T'[s] = min(T[s], T[s ^ upd] + 1), then the closing-row filter, the cap and the lookahead) is one
// pass over the old entries.  Each old entry (k, v) looks up its partner k ^ upd once:
//   key k      (if its closing bits match): old value v against the candidate partner + 1 (partner below the cap);
//   key k^upd  (only when the partner is absent, v below the cap, closing bits match): a new state of value v + 1.
// Every output key arises exactly once (XOR by upd is a bijection and the filter fixes the cleared bits), so the
// outputs need no deduplication: they are appended with one atomic per 1024 old entries and indexed afterwards.
// Counts and least totals are gathered per block before their one atomic: atomics on one cache line serialise at
// about 12 ns each on MI355X, whatever their kind.
#include "tdp_sparse.h"

namespace tdp {

// counters: [0] out count | changed count << 32 (split by k_sp_unpack), [2] overflow flag
__global__ __launch_bounds__(kSpBlock) void k_sp_step(const Ent *__restrict__ old, uint64_t n,
                                                      const uint32_t *__restrict__ slots, uint64_t capmask,
                                                      StepArgs a, BoundArgs b, Ent *__restrict__ out,
                                                      Ent *__restrict__ ch, unsigned long long *counters) {
    sp_step_tile<false>(old, n, (uint64_t)blockIdx.x * kSpBlock * kSpTile, slots, capmask, a, b, out, ch,
                        &counters[0], &counters[2], nullptr, 0);
}

__global__ void k_sp_unpack(unsigned long long *counters) {
    if (threadIdx.x == 0) {
        const unsigned long long p = counters[0];
        counters[0] = p & 0xFFFFFFFFull;
        counters[1] = p >> 32;
    }
}

__global__ __launch_bounds__(kSpBlock) void k_fill_slots(uint32_t *slots, uint64_t cap) {
    uint64_t i = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x;
    if (i < cap) slots[i] = kEmpty;
}

// keys are distinct: an entry takes the first empty slot of its probe sequence
__global__ __launch_bounds__(kSpBlock) void k_index(const Ent *__restrict__ ents, uint64_t n, uint32_t *slots,
                                                    uint64_t capmask) {
    uint64_t i = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x;
    if (i < n) sp_insert(ents, i, slots, capmask);
}

__global__ __launch_bounds__(kSpBlock) void k_lookup(const Ent *__restrict__ ents, const uint32_t *__restrict__ slots,
                                                     uint64_t capmask, const uint64_t *__restrict__ q, uint64_t m,
                                                     uint8_t *__restrict__ vals) {
    uint64_t i = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x;
    if (i < m) vals[i] = (uint8_t)sp_find(ents, slots, capmask, q[2 * i], q[2 * i + 1]);
}

__global__ __launch_bounds__(kSpBlock) void k_truncate(const Ent *__restrict__ ents, uint64_t n, int below,
                                                       Ent *__restrict__ out, unsigned long long *count) {
    const uint64_t base = (uint64_t)blockIdx.x * kSpBlock * kSpTile;
    Ent x[kSpTile];
    bool keep[kSpTile];
    uint64_t pos[kSpTile];
#pragma unroll
    for (int k = 0; k < kSpTile; k++) {
        const uint64_t i = base + (uint64_t)k * kSpBlock + threadIdx.x;
        x[k] = i < n ? ents[i] : Ent{0, 0};
        keep[k] = i < n && (int)(x[k].hv >> 56) < below;
    }
    block_append_tile(keep, count, pos);
#pragma unroll
    for (int k = 0; k < kSpTile; k++)
        if (keep[k]) out[pos[k]] = x[k];
}

// join: best[0] = least total; then among those the least key (hi, then lo), and its forward value
__global__ __launch_bounds__(kSpBlock) void k_join_total(const Ent *__restrict__ f, uint64_t nf,
                                                         const Ent *__restrict__ bents,
                                                         const uint32_t *__restrict__ bslots, uint64_t capmask,
                                                         uint64_t tm_lo, uint64_t tm_hi, unsigned long long *best) {
    // each thread takes entries base + k * kSpBlock; one atomic per block on best[0]
    const uint64_t base = (uint64_t)blockIdx.x * kSpBlock * kSpTile;
    Ent x[kSpTile];
    bool live[kSpTile];
    uint32_t vb[kSpTile];
#pragma unroll
    for (int k = 0; k < kSpTile; k++) {
        const uint64_t i = base + (uint64_t)k * kSpBlock + threadIdx.x;
        live[k] = i < nf;
        x[k] = live[k] ? f[i] : Ent{0, 0};
    }
    sp_find_tile(bents, bslots, capmask, x, live, tm_lo, tm_hi, vb);
    uint32_t t = ~0u;
#pragma unroll
    for (int k = 0; k < kSpTile; k++) {
        const uint32_t s = vb[k] != 255u ? (uint32_t)(x[k].hv >> 56) + vb[k] : ~0u;
        t = s < t ? s : t;
    }
    block_min_atomic(t, &best[0]);
}

__global__ __launch_bounds__(kSpBlock) void k_join_pick(const Ent *__restrict__ f, uint64_t nf,
                                                        const Ent *__restrict__ bents,
                                                        const uint32_t *__restrict__ bslots, uint64_t capmask,
                                                        uint64_t tm_lo, uint64_t tm_hi, int stage,
                                                        unsigned long long *best) {
    uint64_t i = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x;
    if (i >= nf || best[0] == ~0ull) return;          // no join total: nothing to pick
    Ent x = f[i];
    uint64_t hi = x.hv & kKeyMask;
    if (stage >= 2 && hi != best[1]) return;
    if (stage >= 3 && x.lo != best[2]) return;
    uint32_t vb = sp_find(bents, bslots, capmask, x.lo ^ tm_lo, hi ^ tm_hi);
    if (vb == 255u || (x.hv >> 56) + vb != best[0]) return;
    if (stage == 1) atomicMin(&best[1], (unsigned long long)hi);
    else if (stage == 2) atomicMin(&best[2], (unsigned long long)x.lo);
    else best[3] = x.hv >> 56;
}

__global__ __launch_bounds__(kSpBlock) void k_sp_hist(const Ent *__restrict__ ents, uint64_t n,
                                                      unsigned long long *hist) {
    __shared__ uint32_t bins[256];
    bins[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kSpBlock)
        atomicAdd(&bins[ents[i].hv >> 56], 1u);
    __syncthreads();
    if (bins[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)bins[threadIdx.x]);
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v, int n) {
    int i = threadIdx.x;
    if (i < n) p[i] = v;
}

static unsigned sgrid(uint64_t items) {
    uint64_t g = (items + kSpBlock - 1) / kSpBlock;
    return (unsigned)(g ? g : 1);
}

// blocks of the kernels that take kSpTile entries per thread
static unsigned tgrid(uint64_t items) {
    uint64_t g = (items + kSpBlock * kSpTile - 1) / (kSpBlock * kSpTile);
    return (unsigned)(g ? g : 1);
}

}  // namespace tdp

using namespace tdp;

extern "C" {

int tdp_sp_index(const void *d_ents, uint64_t n, uint32_t *d_slots, uint64_t capacity, void *stream) {
    TDP_REQUIRE(capacity && !(capacity & (capacity - 1)) && capacity >= 2 * n && n < 0xFFFFFFFFull,
                "index capacity %llu for %llu entries", (unsigned long long)capacity, (unsigned long long)n);
    hipStream_t st = as_stream(stream);
    k_fill_slots<<<sgrid(capacity), kSpBlock, 0, st>>>(d_slots, capacity);
    TDP_LAUNCH_CHECK();
    if (n) {
        k_index<<<sgrid(n), kSpBlock, 0, st>>>(static_cast<const Ent *>(d_ents), n, d_slots, capacity - 1);
        TDP_LAUNCH_CHECK();
    }
    return TDP_OK;
}

int tdp_sp_step(const void *d_old, uint64_t n, const uint32_t *d_slots, uint64_t capacity, const uint64_t *upd,
                const uint64_t *sel_mask, const uint64_t *sel_bits, int cap, int keep_below, int ncls,
                const uint64_t *d_cls_mask, const int64_t *d_cls_w, int64_t extra, const uint64_t *tmask, int K,
                int use_h, void *d_out, uint64_t out_cap, void *d_ch, uint64_t ch_cap,
                unsigned long long *d_counters, void *stream) {
    TDP_REQUIRE(capacity && !(capacity & (capacity - 1)), "index capacity %llu", (unsigned long long)capacity);
    TDP_REQUIRE(cap >= 0 && cap <= 254, "cap %d outside 0..254", cap);
    TDP_REQUIRE(n < (1ull << 31), "sparse step of %llu states (fewer than 2^31)", (unsigned long long)n);
    hipStream_t st = as_stream(stream);
    StepArgs a;
    a.upd_lo = upd[0];
    a.upd_hi = upd[1] & kKeyMask;
    a.sm_lo = sel_mask[0];
    a.sm_hi = sel_mask[1] & kKeyMask;
    a.sb_lo = sel_bits[0];
    a.sb_hi = sel_bits[1] & kKeyMask;
    a.cap = cap;
    int kb = keep_below > 0 ? keep_below : (1 << 30);
    a.kb = cap + 1 < kb ? cap + 1 : kb;
    a.want_changed = d_ch != nullptr;
    a.out_cap = out_cap;
    a.ch_cap = ch_cap;
    BoundArgs b;
    b.mask = d_cls_mask;
    b.w = d_cls_w;
    b.ncls = use_h ? ncls : 0;
    b.on = use_h;
    b.extra = extra;
    b.tm_lo = tmask[0];
    b.tm_hi = tmask[1] & kKeyMask;
    b.K = K;
    k_set_u64<<<1, 64, 0, st>>>(d_counters, 0ull, 3);
    TDP_LAUNCH_CHECK();
    if (n) {
        k_sp_step<<<tgrid(n), kSpBlock, 0, st>>>(static_cast<const Ent *>(d_old), n, d_slots, capacity - 1, a, b,
                                                 static_cast<Ent *>(d_out), static_cast<Ent *>(d_ch), d_counters);
        TDP_LAUNCH_CHECK();
        k_sp_unpack<<<1, 64, 0, st>>>(d_counters);
        TDP_LAUNCH_CHECK();
    }
    return TDP_OK;
}

int tdp_sp_lookup(const void *d_ents, const uint32_t *d_slots, uint64_t capacity, const uint64_t *d_keys, uint64_t m,
                  uint8_t *d_vals, void *stream) {
    if (!m) return TDP_OK;
    hipStream_t st = as_stream(stream);
    k_lookup<<<sgrid(m), kSpBlock, 0, st>>>(static_cast<const Ent *>(d_ents), d_slots, capacity - 1, d_keys, m,
                                            d_vals);
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

int tdp_sp_truncate(const void *d_ents, uint64_t n, int below, void *d_out, unsigned long long *d_count,
                    void *stream) {
    hipStream_t st = as_stream(stream);
    k_set_u64<<<1, 64, 0, st>>>(d_count, 0ull, 1);
    TDP_LAUNCH_CHECK();
    if (n) {
        k_truncate<<<tgrid(n), kSpBlock, 0, st>>>(static_cast<const Ent *>(d_ents), n, below,
                                                  static_cast<Ent *>(d_out), d_count);
        TDP_LAUNCH_CHECK();
    }
    return TDP_OK;
}

int tdp_sp_join(const void *d_f, uint64_t nf, const void *d_bents, const uint32_t *d_bslots, uint64_t bcapacity,
                const uint64_t *tmask, unsigned long long *d_best, void *stream) {
    hipStream_t st = as_stream(stream);
    k_set_u64<<<1, 64, 0, st>>>(d_best, ~0ull, 4);
    TDP_LAUNCH_CHECK();
    if (!nf) return TDP_OK;
    const Ent *f = static_cast<const Ent *>(d_f);
    const Ent *be = static_cast<const Ent *>(d_bents);
    uint64_t tlo = tmask[0], thi = tmask[1] & kKeyMask;
    k_join_total<<<tgrid(nf), kSpBlock, 0, st>>>(f, nf, be, d_bslots, bcapacity - 1, tlo, thi, d_best);
    TDP_LAUNCH_CHECK();
    for (int stage = 1; stage <= 3; stage++) {
        k_join_pick<<<sgrid(nf), kSpBlock, 0, st>>>(f, nf, be, d_bslots, bcapacity - 1, tlo, thi, stage, d_best);
        TDP_LAUNCH_CHECK();
    }
    return TDP_OK;
}

int tdp_sp_hist(const void *d_ents, uint64_t n, uint64_t *d_hist, void *stream) {
    hipStream_t st = as_stream(stream);
    unsigned long long *h = reinterpret_cast<unsigned long long *>(d_hist);
    k_set_u64<<<1, 256, 0, st>>>(h, 0ull, 256);
    TDP_LAUNCH_CHECK();
    if (n) {
        uint64_t g = sgrid(n);
        k_sp_hist<<<(unsigned)(g > 2048 ? 2048 : g), kSpBlock, 0, st>>>(static_cast<const Ent *>(d_ents), n, h);
        TDP_LAUNCH_CHECK();
    }
    return TDP_OK;
}

}  // extern "C"
