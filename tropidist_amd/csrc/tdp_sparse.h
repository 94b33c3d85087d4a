// Device pieces shared by the sparse table kernels (tdp_sparse.hip: one table per call; tdp_sparse_batch.hip: many
// sweeps per launch).  Table layout, hashing, probing and the per-entry column-step rule of
// cdist/sparse_dp.c:116-202 live here once.
#pragma once
#include "tdp_common.h"

namespace tdp {

constexpr int kSpBlock = 256;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint64_t kKeyMask = (1ull << 56) - 1;

struct Ent {
    uint64_t lo, hv;
};

struct Key128 {
    uint64_t lo, hi;
};

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

__device__ __forceinline__ uint64_t hash_key(uint64_t lo, uint64_t hi) { return mix64(lo ^ mix64(hi + 0x9e3779b97f4a7c15ull)); }

// value of key (lo, hi) in the table, 255 if absent
__device__ __forceinline__ uint32_t sp_find(const Ent *__restrict__ ents, const uint32_t *__restrict__ slots,
                                            uint64_t capmask, uint64_t lo, uint64_t hi) {
    uint64_t h = hash_key(lo, hi) & capmask;
    for (;;) {
        uint32_t e = slots[h];
        if (e == kEmpty) return 255u;
        Ent x = ents[e];
        if (x.lo == lo && (x.hv & kKeyMask) == hi) return (uint32_t)(x.hv >> 56);
        h = (h + 1) & capmask;
    }
}

// Lookahead bound (cdist/sparse_dp.c:25-44): ceil((extra + sum_c w_c |(key ^ tmask) & mask_c|) / Q).
struct BoundArgs {
    const uint64_t *mask;     // 2 words per class (lo, hi)
    const int64_t *w;
    int ncls;
    int on;
    int64_t extra;
    uint64_t tm_lo, tm_hi;
    int K;
};

__device__ __forceinline__ bool keep_state(const BoundArgs &b, uint64_t lo, uint64_t hi, int v) {
    if (!b.on) return true;
    uint64_t ulo = lo ^ b.tm_lo, uhi = hi ^ b.tm_hi;
    int64_t tot = b.extra;
    for (int c = 0; c < b.ncls; c++)
        tot += b.w[c] * (int64_t)(__popcll(ulo & b.mask[2 * c]) + __popcll(uhi & b.mask[2 * c + 1]));
    const int64_t Q = 720720;
    return v + (int)((tot + Q - 1) / Q) <= b.K;
}

// Wave-aggregated append: returns the slot of this lane's item in `out` (only meaningful where pred), one atomic
// per wave.  Must be reached by every lane of the wave.
__device__ __forceinline__ uint64_t wave_append(unsigned long long *counter, bool pred) {
    uint64_t ballot = __ballot(pred);
    uint32_t lane = __lane_id();
    uint64_t below = ballot & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
    uint32_t rank = (uint32_t)__popcll(below);
    uint32_t total = (uint32_t)__popcll(ballot);
    int leader = ballot ? __ffsll((long long)ballot) - 1 : 0;
    unsigned long long base = 0;
    if (pred && (int)lane == leader) base = atomicAdd(counter, (unsigned long long)total);
    base = __shfl(base, leader, 64);
    return base + rank;
}

// Least value over the 64 lanes of the wave (every lane must call it).
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const uint32_t w = (uint32_t)__shfl_xor((int)v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

struct StepArgs {
    uint64_t upd_lo, upd_hi;
    uint64_t sm_lo, sm_hi, sb_lo, sb_hi;      // closing rows: mask and required bits
    int cap, kb;                              // candidates from values < cap; table keeps values < kb
    int want_changed;
    uint64_t out_cap, ch_cap;
};

// Inserts entry i of a table into its index (keys are distinct: the entry takes the first empty slot of its probe
// sequence; the index has at least one empty slot, so the probe ends).
__device__ __forceinline__ void sp_insert(const Ent *__restrict__ ents, uint64_t i, uint32_t *slots,
                                          uint64_t capmask) {
    Ent x = ents[i];
    uint64_t h = hash_key(x.lo, x.hv & kKeyMask) & capmask;
    while (atomicCAS(&slots[h], kEmpty, (uint32_t)i) != kEmpty) h = (h + 1) & capmask;
}

// One old entry's contribution to a column step (T'[s] = min(T[s], T[s ^ upd] + 1), the closing-row filter, the cap
// and the lookahead, cdist/sparse_dp.c:118-202): up to two table entries (o1/e1, o2/e2) and two changed entries
// (c1/f1, c2/f2).  Entry i of `old` with i >= n contributes nothing.
struct StepOut {
    bool o1, o2, c1, c2;
    Ent e1, e2, f1, f2;
};

__device__ __forceinline__ StepOut sp_step_entry(const Ent *__restrict__ old, uint64_t i, uint64_t n,
                                                 const uint32_t *__restrict__ slots, uint64_t capmask,
                                                 const StepArgs &a, const BoundArgs &b) {
    StepOut r;
    r.o1 = r.o2 = r.c1 = r.c2 = false;
    r.e1 = r.e2 = r.f1 = r.f2 = Ent{0, 0};
    if (i >= n) return r;
    Ent x = old[i];
    uint64_t klo = x.lo, khi = x.hv & kKeyMask;
    int v = (int)(x.hv >> 56);
    uint64_t qlo = klo ^ a.upd_lo, qhi = khi ^ a.upd_hi;
    bool has_upd = (a.upd_lo | a.upd_hi) != 0;
    uint32_t vq = has_upd ? sp_find(old, slots, capmask, qlo, qhi) : 255u;
    uint64_t clo = ~a.sm_lo, chi = ~a.sm_hi & kKeyMask;
    // key k
    if ((klo & a.sm_lo) == a.sb_lo && (khi & a.sm_hi) == a.sb_hi) {
        uint64_t alo = klo & clo, ahi = khi & chi;
        bool cand = has_upd && vq != 255u && (int)vq < a.cap && keep_state(b, alo, ahi, (int)vq + 1);
        int vc = (int)vq + 1;
        if (cand && vc < v) {
            Ent c = {alo, ahi | ((uint64_t)vc << 56)};
            if (a.want_changed) { r.c1 = true; r.f1 = c; }
            if (vc < a.kb) { r.o1 = true; r.e1 = c; }
        } else if (v < a.kb && keep_state(b, alo, ahi, v)) {
            r.o1 = true;
            r.e1 = Ent{alo, ahi | ((uint64_t)v << 56)};
        }
    }
    // key k ^ upd, a new state when the partner is absent
    if (has_upd && vq == 255u && v < a.cap && (qlo & a.sm_lo) == a.sb_lo && (qhi & a.sm_hi) == a.sb_hi) {
        uint64_t alo = qlo & clo, ahi = qhi & chi;
        if (keep_state(b, alo, ahi, v + 1)) {
            Ent c = {alo, ahi | ((uint64_t)(v + 1) << 56)};
            if (a.want_changed) { r.c2 = true; r.f2 = c; }
            if (v + 1 < a.kb) { r.o2 = true; r.e2 = c; }
        }
    }
    return r;
}

// Appends a step's outputs with one atomic per block and counter: every thread of the block must call it, the same
// number of times (it holds three block barriers).  Each wave counts its outputs by ballot, thread 0 and 1 reserve
// the block's table and changed ranges, and each lane writes at its wave's offset plus its rank.  One reservation
// per block instead of four per wave keeps the two counters, which every block of the launch shares, from
// serialising the step.  counters: [0] table count, [1] changed count, [2] overflow flag (an output past out_cap /
// ch_cap).
__device__ __forceinline__ void sp_step_emit(const StepOut &r, Ent *__restrict__ out, uint64_t out_cap,
                                             Ent *__restrict__ ch, uint64_t ch_cap, unsigned long long *counters) {
    constexpr int kWaves = kSpBlock / 64;
    __shared__ uint32_t s_cnt[2][kWaves];
    __shared__ unsigned long long s_base[2];
    const uint32_t lane = __lane_id();
    const uint32_t w = threadIdx.x >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint64_t bo1 = __ballot(r.o1), bo2 = __ballot(r.o2), bc1 = __ballot(r.c1), bc2 = __ballot(r.c2);
    const uint32_t no1 = (uint32_t)__popcll(bo1), nc1 = (uint32_t)__popcll(bc1);
    if (lane == 0) {
        s_cnt[0][w] = no1 + (uint32_t)__popcll(bo2);
        s_cnt[1][w] = nc1 + (uint32_t)__popcll(bc2);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t tot = 0;
        for (int k = 0; k < kWaves; k++) tot += s_cnt[threadIdx.x][k];
        s_base[threadIdx.x] = tot ? atomicAdd(&counters[threadIdx.x], (unsigned long long)tot) : 0ull;
    }
    __syncthreads();
    uint64_t bo = s_base[0], bc = s_base[1];
    for (uint32_t k = 0; k < w; k++) {
        bo += s_cnt[0][k];
        bc += s_cnt[1][k];
    }
    __syncthreads();                          // the next call rewrites s_cnt and s_base
    if (r.o1) {
        const uint64_t p = bo + (uint64_t)__popcll(bo1 & lt);
        if (p < out_cap) out[p] = r.e1; else counters[2] = 1;
    }
    if (r.o2) {
        const uint64_t p = bo + no1 + (uint64_t)__popcll(bo2 & lt);
        if (p < out_cap) out[p] = r.e2; else counters[2] = 1;
    }
    if (r.c1) {
        const uint64_t p = bc + (uint64_t)__popcll(bc1 & lt);
        if (p < ch_cap) ch[p] = r.f1; else counters[2] = 1;
    }
    if (r.c2) {
        const uint64_t p = bc + nc1 + (uint64_t)__popcll(bc2 & lt);
        if (p < ch_cap) ch[p] = r.f2; else counters[2] = 1;
    }
}

}  // namespace tdp
