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

// value of key (lo, hi) in the table, 255 if absent, probing from slot h (the key's hash slot unless continuing)
__device__ __forceinline__ uint32_t sp_find(const Ent *__restrict__ ents, const uint32_t *__restrict__ slots,
                                            uint64_t capmask, uint64_t lo, uint64_t hi, uint64_t h) {
    for (;;) {
        uint32_t e = slots[h];
        if (e == kEmpty) return 255u;
        Ent x = ents[e];
        if (x.lo == lo && (x.hv & kKeyMask) == hi) return (uint32_t)(x.hv >> 56);
        h = (h + 1) & capmask;
    }
}

__device__ __forceinline__ uint32_t sp_find(const Ent *__restrict__ ents, const uint32_t *__restrict__ slots,
                                            uint64_t capmask, uint64_t lo, uint64_t hi) {
    return sp_find(ents, slots, capmask, lo, hi, hash_key(lo, hi) & capmask);
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
// and the lookahead, cdist/sparse_dp.c:118-202), given the value vq of its partner key ^ upd in the old table (255
// if absent or if the column has no rows): up to two table entries and two changed entries, as flags.
//   kO1 / kC1: key & ~sel_mask, as a table / changed entry, with the partner's value + 1 if kImp, else its own value
//   kO2 / kC2: (key ^ upd) & ~sel_mask with its value + 1 (a new state: the partner was absent)
constexpr uint32_t kO1 = 1, kO2 = 2, kC1 = 4, kC2 = 8, kImp = 16;

__device__ __forceinline__ uint32_t sp_flags(const Ent &x, uint32_t vq, const StepArgs &a, const BoundArgs &b) {
    const uint64_t klo = x.lo, khi = x.hv & kKeyMask;
    const int v = (int)(x.hv >> 56);
    const bool has_upd = (a.upd_lo | a.upd_hi) != 0;
    const uint64_t clo = ~a.sm_lo, chi = ~a.sm_hi & kKeyMask;
    uint32_t f = 0;
    // key k
    if ((klo & a.sm_lo) == a.sb_lo && (khi & a.sm_hi) == a.sb_hi) {
        const uint64_t alo = klo & clo, ahi = khi & chi;
        const int vc = (int)vq + 1;
        const bool cand = has_upd && vq != 255u && (int)vq < a.cap && keep_state(b, alo, ahi, vc);
        if (cand && vc < v) {
            f |= kImp;
            if (a.want_changed) f |= kC1;
            if (vc < a.kb) f |= kO1;
        } else if (v < a.kb && keep_state(b, alo, ahi, v)) {
            f |= kO1;
        }
    }
    // key k ^ upd, a new state when the partner is absent
    const uint64_t qlo = klo ^ a.upd_lo, qhi = khi ^ a.upd_hi;
    if (has_upd && vq == 255u && v < a.cap && (qlo & a.sm_lo) == a.sb_lo && (qhi & a.sm_hi) == a.sb_hi &&
        keep_state(b, qlo & clo, qhi & chi, v + 1)) {
        if (a.want_changed) f |= kC2;
        if (v + 1 < a.kb) f |= kO2;
    }
    return f;
}

__device__ __forceinline__ Ent sp_first(const Ent &x, uint32_t vq, uint32_t f, const StepArgs &a) {
    const uint64_t val = (f & kImp) ? (uint64_t)vq + 1 : x.hv >> 56;
    return Ent{x.lo & ~a.sm_lo, (x.hv & kKeyMask & ~a.sm_hi) | (val << 56)};
}

__device__ __forceinline__ Ent sp_second(const Ent &x, const StepArgs &a) {
    return Ent{(x.lo ^ a.upd_lo) & ~a.sm_lo,
               (((x.hv & kKeyMask) ^ a.upd_hi) & ~a.sm_hi & kKeyMask) | (((x.hv >> 56) + 1) << 56)};
}

constexpr int kSpTile = 4;            // entries per thread of one step tile (a block steps 1024 entries)

// Partner values of kSpTile entries at once: every probe sequence's first slot and first entry are loaded together,
// and only probes that meet another key continue one by one.
__device__ __forceinline__ void sp_find_tile(const Ent *__restrict__ ents, const uint32_t *__restrict__ slots,
                                             uint64_t capmask, const Ent (&x)[kSpTile], const bool (&live)[kSpTile],
                                             uint64_t upd_lo, uint64_t upd_hi, uint32_t (&vq)[kSpTile]) {
    uint64_t h[kSpTile];
    uint32_t e[kSpTile];
#pragma unroll
    for (int k = 0; k < kSpTile; k++) {
        h[k] = hash_key(x[k].lo ^ upd_lo, (x[k].hv & kKeyMask) ^ upd_hi) & capmask;
        e[k] = live[k] ? slots[h[k]] : kEmpty;
    }
#pragma unroll
    for (int k = 0; k < kSpTile; k++) {
        vq[k] = 255u;
        if (e[k] == kEmpty) continue;
        const uint64_t qlo = x[k].lo ^ upd_lo, qhi = (x[k].hv & kKeyMask) ^ upd_hi;
        const Ent y = ents[e[k]];
        if (y.lo == qlo && (y.hv & kKeyMask) == qhi) vq[k] = (uint32_t)(y.hv >> 56);
        else vq[k] = sp_find(ents, slots, capmask, qlo, qhi, (h[k] + 1) & capmask);
    }
}

// Steps the entries [base, base + kSpTile * kSpBlock) of `old` (entry base + e * kSpBlock + threadIdx.x in round e)
// and appends their outputs with one reservation for the whole block: a 64-bit atomic add on *pack, whose low half
// counts table entries and high half changed entries (every table holds fewer than 2^31 entries).  Atomics on one
// cache line are serialised by the memory system, about 12 ns each on MI355X, so one per 1024 entries keeps them
// far below the step's memory time.  Each wave counts its outputs per round by ballot; the block's prefix gives every
// wave its offsets.  *over = 1 if an output falls past out_cap / ch_cap.  With INDEX, each table entry written at
// position p is also placed in the index nslots (keys are distinct; the entry takes the first empty slot of its
// probe sequence): the first compare-and-swap of every output is issued before any result is awaited, and only
// the outputs whose home slot was taken probe on.  Every thread of the block must call it (it holds block barriers).
template <bool INDEX>
__device__ __forceinline__ void sp_step_tile(const Ent *__restrict__ old, uint64_t n, uint64_t base,
                                             const uint32_t *__restrict__ slots, uint64_t capmask, const StepArgs &a,
                                             const BoundArgs &b, Ent *__restrict__ out, Ent *__restrict__ ch,
                                             unsigned long long *pack, unsigned long long *over,
                                             uint32_t *__restrict__ nslots, uint64_t ncapmask) {
    constexpr int kWaves = kSpBlock / 64;
    __shared__ uint32_t s_cnt[kSpTile][kWaves];       // table count | changed count << 16, per round and wave
    __shared__ unsigned long long s_base;
    const uint32_t lane = __lane_id();
    const uint32_t w = threadIdx.x >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    Ent x[kSpTile];
    bool live[kSpTile];
    uint32_t vq[kSpTile];
#pragma unroll
    for (int k = 0; k < kSpTile; k++) {
        const uint64_t i = base + (uint64_t)k * kSpBlock + threadIdx.x;
        live[k] = i < n;
        x[k] = live[k] ? old[i] : Ent{0, 0};
    }
    const bool has_upd = (a.upd_lo | a.upd_hi) != 0;
    if (has_upd) {
        sp_find_tile(old, slots, capmask, x, live, a.upd_lo, a.upd_hi, vq);
    } else {
#pragma unroll
        for (int k = 0; k < kSpTile; k++) vq[k] = 255u;
    }
    uint32_t fl[kSpTile];
#pragma unroll
    for (int k = 0; k < kSpTile; k++) {
        fl[k] = live[k] ? sp_flags(x[k], vq[k], a, b) : 0u;
        const uint32_t no = (uint32_t)__popcll(__ballot(fl[k] & kO1)) + (uint32_t)__popcll(__ballot(fl[k] & kO2));
        const uint32_t nc = (uint32_t)__popcll(__ballot(fl[k] & kC1)) + (uint32_t)__popcll(__ballot(fl[k] & kC2));
        if (lane == 0) s_cnt[k][w] = no | (nc << 16);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int k = 0; k < kSpTile; k++)
            for (int q = 0; q < kWaves; q++)
                tot += (s_cnt[k][q] & 0xFFFFu) | ((unsigned long long)(s_cnt[k][q] >> 16) << 32);
        s_base = tot ? atomicAdd(pack, tot) : 0ull;
    }
    __syncthreads();
    uint64_t bo = s_base & 0xFFFFFFFFull, bc = s_base >> 32;
    uint32_t got[2 * kSpTile], posv[2 * kSpTile];
#pragma unroll
    for (int k = 0; k < 2 * kSpTile; k++) posv[k] = kEmpty;
#pragma unroll
    for (int k = 0; k < kSpTile; k++) {
        uint64_t wo = bo, wc = bc;
        for (int q = 0; q < kWaves; q++) {
            const uint32_t c = s_cnt[k][q];
            if (q < (int)w) {
                wo += c & 0xFFFFu;
                wc += c >> 16;
            }
            bo += c & 0xFFFFu;
            bc += c >> 16;
        }
        const uint32_t f = fl[k];
        const uint64_t bo1 = __ballot(f & kO1), bo2 = __ballot(f & kO2);
        const uint64_t bc1 = __ballot(f & kC1), bc2 = __ballot(f & kC2);
        if (f & (kO1 | kC1)) {
            const Ent e1 = sp_first(x[k], vq[k], f, a);
            if (f & kO1) {
                const uint64_t p = wo + (uint64_t)__popcll(bo1 & lt);
                if (p < a.out_cap) {
                    out[p] = e1;
                    if (INDEX) {
                        posv[2 * k] = (uint32_t)p;
                        got[2 * k] = atomicCAS(&nslots[hash_key(e1.lo, e1.hv & kKeyMask) & ncapmask], kEmpty,
                                               (uint32_t)p);
                    }
                } else {
                    *over = 1;
                }
            }
            if (f & kC1) {
                const uint64_t p = wc + (uint64_t)__popcll(bc1 & lt);
                if (p < a.ch_cap) ch[p] = e1; else *over = 1;
            }
        }
        if (f & (kO2 | kC2)) {
            const Ent e2 = sp_second(x[k], a);
            if (f & kO2) {
                const uint64_t p = wo + (uint64_t)__popcll(bo1) + (uint64_t)__popcll(bo2 & lt);
                if (p < a.out_cap) {
                    out[p] = e2;
                    if (INDEX) {
                        posv[2 * k + 1] = (uint32_t)p;
                        got[2 * k + 1] = atomicCAS(&nslots[hash_key(e2.lo, e2.hv & kKeyMask) & ncapmask], kEmpty,
                                                   (uint32_t)p);
                    }
                } else {
                    *over = 1;
                }
            }
            if (f & kC2) {
                const uint64_t p = wc + (uint64_t)__popcll(bc1) + (uint64_t)__popcll(bc2 & lt);
                if (p < a.ch_cap) ch[p] = e2; else *over = 1;
            }
        }
    }
    if (INDEX) {
#pragma unroll
        for (int k = 0; k < 2 * kSpTile; k++) {
            if (posv[k] == kEmpty || got[k] == kEmpty) continue;
            const Ent e = (k & 1) ? sp_second(x[k >> 1], a) : sp_first(x[k >> 1], vq[k >> 1], fl[k >> 1], a);
            uint64_t h = ((hash_key(e.lo, e.hv & kKeyMask) & ncapmask) + 1) & ncapmask;
            while (atomicCAS(&nslots[h], kEmpty, posv[k]) != kEmpty) h = (h + 1) & ncapmask;
        }
    }
    __syncthreads();                          // the next call rewrites s_cnt and s_base
}

// The least of per-thread values over the block, one atomicMin on *dst from thread 0 (every thread must call it).
__device__ __forceinline__ void block_min_atomic(uint32_t t, unsigned long long *dst) {
    __shared__ uint32_t s_min[kSpBlock / 64];
    t = wave_min_u32(t);
    if (__lane_id() == 0) s_min[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = s_min[0];
        for (int q = 1; q < kSpBlock / 64; q++) m = s_min[q] < m ? s_min[q] : m;
        // a block that cannot lower the current least (read without an atomic; a stale read only costs one more
        // atomic) skips it
        if (m != ~0u && (unsigned long long)m < *reinterpret_cast<volatile unsigned long long *>(dst))
            atomicMin(dst, (unsigned long long)m);
    }
    __syncthreads();
}

// Compacting append of up to kSpTile items per thread (item k of this thread kept where keep[k]): returns through
// pos[k] each kept item's position, with one atomic add on *counter for the block.  Every thread must call it.
__device__ __forceinline__ void block_append_tile(const bool (&keep)[kSpTile], unsigned long long *counter,
                                                  uint64_t (&pos)[kSpTile]) {
    constexpr int kWaves = kSpBlock / 64;
    __shared__ uint32_t s_n[kSpTile][kWaves];
    __shared__ unsigned long long s_at;
    const uint32_t lane = __lane_id();
    const uint32_t w = threadIdx.x >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint64_t bal[kSpTile];
#pragma unroll
    for (int k = 0; k < kSpTile; k++) {
        bal[k] = __ballot(keep[k]);
        if (lane == 0) s_n[k][w] = (uint32_t)__popcll(bal[k]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int k = 0; k < kSpTile; k++)
            for (int q = 0; q < kWaves; q++) tot += s_n[k][q];
        s_at = tot ? atomicAdd(counter, tot) : 0ull;
    }
    __syncthreads();
    uint64_t at = s_at;
#pragma unroll
    for (int k = 0; k < kSpTile; k++) {
        uint64_t mine = at;
        for (int q = 0; q < kWaves; q++) {
            if (q < (int)w) mine += s_n[k][q];
            at += s_n[k][q];
        }
        pos[k] = mine + (uint64_t)__popcll(bal[k] & lt);
    }
    __syncthreads();
}

}  // namespace tdp
