// Dense tables of the sparse sweeps' fallback on gfx950 (the Dense table and its column step of
// cdist/sparse_dp.py:221-289 and 680-707).
//
// A dense table holds every state of the active rows: T[i], uint8 (255 = absent), for the compact index i of a state
// (bit k of i = the state's bit at slot slots[k], slots ascending), 2^w entries.  A sweep switches to it where a
// sparse table fills more than a given fraction of its 2^w states and back where it empties; the tables hold exactly
// the states and values the sparse tables would, so values, counts and witnesses do not change.
//
// One column step is one gather pass writing the new table: output entry j is placed in the "mid" space (the old
// slots and the rows starting at the column) by inserting the ending rows' target bits, the old values of it and of
// its partner (xor the column's rows) are read -- absent where a starting row's bit is set -- and
// min(T[x], T[x ^ cm] + 1) is kept if it is within the cap and the lookahead bound.  Every output entry is written
// once, with no atomics but one count per wave.
#include "tdp_sparse.h"

namespace tdp {

constexpr int kDnMaxBits = 40;        // the widest dense table (2^40 bytes)
constexpr int kDnMaxCls = 64;         // lookahead classes of one cut with a nonzero compact mask
constexpr uint32_t kDnGridCap = 1u << 20;

struct DnStepArgs {
    uint64_t cm;                      // the column's rows, mid space
    uint64_t start_mask;              // the starting rows' bits, mid space
    uint64_t tmask;                   // target row, new compact space
    int64_t extra;
    int32_t nstart, nend, ncls, cap, K, use_h;
    uint8_t start_pos[kDnMaxBits + 8];        // descending mid positions of the starting rows
    uint8_t end_pos[kDnMaxBits + 8];          // ascending mid positions of the ending rows
    uint8_t end_bit[kDnMaxBits + 8];          // their targets
    uint64_t cls_mask[kDnMaxCls];             // new compact space
    int64_t cls_w[kDnMaxCls];
};

struct DnSlots {                      // the compact layout of one table
    uint64_t act_lo, act_hi;          // the active slots as a 120-bit mask
    int32_t w, pad;
    uint8_t slot[kDnMaxBits + 8];     // ascending
};

__device__ __forceinline__ uint64_t insert_bit(uint64_t x, int pos, uint64_t b) {
    const uint64_t low = x & ((1ull << pos) - 1);
    return ((x >> pos) << (pos + 1)) | (b << pos) | low;
}

__device__ __forceinline__ uint64_t remove_bit(uint64_t x, int pos) {
    const uint64_t low = x & ((1ull << pos) - 1);
    return ((x >> (pos + 1)) << pos) | low;
}

// compact index of a state whose bits lie in the active slots
__device__ __forceinline__ uint64_t dn_compact(const DnSlots &s, uint64_t lo, uint64_t hi) {
    uint64_t i = 0;
    for (int k = 0; k < s.w; k++) {
        const int b = s.slot[k];
        const uint64_t bit = b < 64 ? (lo >> b) & 1ull : (hi >> (b - 64)) & 1ull;
        i |= bit << k;
    }
    return i;
}

__device__ __forceinline__ void dn_expand(const DnSlots &s, uint64_t i, uint64_t &lo, uint64_t &hi) {
    lo = hi = 0;
    for (int k = 0; k < s.w; k++) {
        const uint64_t bit = (i >> k) & 1ull;
        const int b = s.slot[k];
        if (b < 64) lo |= bit << b;
        else hi |= bit << (b - 64);
    }
}

// value of a state in a dense table (255 if absent or outside the active slots)
__device__ __forceinline__ uint32_t dn_find(const uint8_t *__restrict__ T, const DnSlots &s, uint64_t lo,
                                            uint64_t hi) {
    if ((lo & ~s.act_lo) | (hi & ~s.act_hi)) return 255u;
    return T[dn_compact(s, lo, hi)];
}

__device__ __forceinline__ uint32_t dn_mid(const uint8_t *__restrict__ T, const DnStepArgs &a, uint64_t x) {
    if (x & a.start_mask) return 255u;
    for (int k = 0; k < a.nstart; k++) x = remove_bit(x, a.start_pos[k]);
    return T[x];
}

// ---------------------------------------------------------------------------------------------- kernels
__global__ __launch_bounds__(kSpBlock) void k_dn_fill(uint32_t *__restrict__ T4, uint64_t nwords) {
    const uint64_t stride = (uint64_t)gridDim.x * kSpBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x; i < nwords; i += stride) T4[i] = 0xFFFFFFFFu;
}

__global__ void k_dn_fill_tail(uint8_t *T, uint64_t n) {
    if (threadIdx.x < n) T[threadIdx.x] = 255;
}

__global__ __launch_bounds__(kSpBlock) void k_dn_scatter(const Ent *__restrict__ ents, uint64_t n, DnSlots s,
                                                         uint8_t *__restrict__ T) {
    const uint64_t stride = (uint64_t)gridDim.x * kSpBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x; i < n; i += stride) {
        const Ent x = ents[i];
        T[dn_compact(s, x.lo, x.hv & kKeyMask)] = (uint8_t)(x.hv >> 56);
    }
}

// Loop bounds are block-uniform, so every lane reaches the wave-wide append.
__global__ __launch_bounds__(kSpBlock) void k_dn_gather(const uint8_t *__restrict__ T, uint64_t n, DnSlots s,
                                                        Ent *__restrict__ out, uint64_t cap,
                                                        unsigned long long *counters) {
    const uint64_t stride = (uint64_t)gridDim.x * kSpBlock;
    for (uint64_t base = (uint64_t)blockIdx.x * kSpBlock; base < n; base += stride) {
        const uint64_t i = base + threadIdx.x;
        const uint32_t v = i < n ? T[i] : 255u;
        const bool keep = v != 255u;
        const uint64_t p = wave_append(&counters[0], keep);
        if (keep) {
            uint64_t lo, hi;
            dn_expand(s, i, lo, hi);
            if (p < cap) out[p] = Ent{lo, hi | ((uint64_t)v << 56)};
            else counters[1] = 1;
        }
    }
}

__global__ __launch_bounds__(kSpBlock) void k_dn_step(const uint8_t *__restrict__ T, uint8_t *__restrict__ U,
                                                      uint64_t nout, DnStepArgs a, unsigned long long *count) {
    const uint64_t stride = (uint64_t)gridDim.x * kSpBlock;
    uint32_t kept = 0;
    for (uint64_t j = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x; j < nout; j += stride) {
        uint64_t x = j;
        for (int k = 0; k < a.nend; k++) x = insert_bit(x, a.end_pos[k], a.end_bit[k]);
        uint32_t v = dn_mid(T, a, x);
        if (a.cm) {
            const uint32_t b = dn_mid(T, a, x ^ a.cm);
            const uint32_t b1 = b >= 254u ? 255u : b + 1;
            v = b1 < v ? b1 : v;
        }
        if (v != 255u && (int)v > a.cap) v = 255u;
        if (v != 255u && a.use_h) {
            const uint64_t u = j ^ a.tmask;
            int64_t tot = a.extra;
            for (int c = 0; c < a.ncls; c++) tot += a.cls_w[c] * (int64_t)__popcll(u & a.cls_mask[c]);
            const int64_t Q = 720720;
            if ((int)v + (int)((tot + Q - 1) / Q) > a.K) v = 255u;
        }
        U[j] = (uint8_t)v;
        kept += v != 255u;
    }
    // one atomic per wave
    for (int o = 32; o >= 1; o >>= 1) kept += (uint32_t)__shfl_xor((int)kept, o, 64);
    if (__lane_id() == 0 && kept) atomicAdd(count, (unsigned long long)kept);
}

__global__ __launch_bounds__(kSpBlock) void k_dn_join_total(const Ent *__restrict__ f, uint64_t nf,
                                                            const uint8_t *__restrict__ T, DnSlots s, uint64_t tm_lo,
                                                            uint64_t tm_hi, unsigned long long *best) {
    const uint64_t stride = (uint64_t)gridDim.x * kSpBlock;
    for (uint64_t base = (uint64_t)blockIdx.x * kSpBlock; base < nf; base += stride) {
        const uint64_t i = base + threadIdx.x;
        uint32_t t = ~0u;
        if (i < nf) {
            const Ent x = f[i];
            const uint32_t vb = dn_find(T, s, x.lo ^ tm_lo, (x.hv & kKeyMask) ^ tm_hi);
            if (vb != 255u) t = (uint32_t)(x.hv >> 56) + vb;
        }
        t = wave_min_u32(t);
        if (__lane_id() == 0 && t != ~0u) atomicMin(&best[0], (unsigned long long)t);
    }
}

// stage 1: the least key hi among states reaching the least total; 2: then the least key lo; 3: its value
__global__ __launch_bounds__(kSpBlock) void k_dn_join_pick(const Ent *__restrict__ f, uint64_t nf,
                                                           const uint8_t *__restrict__ T, DnSlots s, uint64_t tm_lo,
                                                           uint64_t tm_hi, int stage, unsigned long long *best) {
    const uint64_t stride = (uint64_t)gridDim.x * kSpBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x; i < nf; i += stride) {
        const Ent x = f[i];
        const uint64_t hi = x.hv & kKeyMask;
        if (stage >= 2 && hi != best[1]) continue;
        if (stage >= 3 && x.lo != best[2]) continue;
        const uint32_t vb = dn_find(T, s, x.lo ^ tm_lo, hi ^ tm_hi);
        if (vb == 255u || (x.hv >> 56) + vb != best[0]) continue;
        if (stage == 1) atomicMin(&best[1], (unsigned long long)hi);
        else if (stage == 2) atomicMin(&best[2], (unsigned long long)x.lo);
        else best[3] = x.hv >> 56;
    }
}

__global__ void k_dn_set(unsigned long long *p, unsigned long long v, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

static unsigned dgrid(uint64_t items) {
    const uint64_t g = (items + kSpBlock - 1) / kSpBlock;
    return (unsigned)(g == 0 ? 1 : (g > kDnGridCap ? kDnGridCap : g));
}

static int dn_slots(const uint8_t *slots, int w, DnSlots &s) {
    TDP_REQUIRE(w >= 0 && w <= kDnMaxBits, "dense table of %d bits (at most %d)", w, kDnMaxBits);
    memset(&s, 0, sizeof(s));
    s.w = w;
    for (int k = 0; k < w; k++) {
        TDP_REQUIRE(slots[k] < 120 && (k == 0 || slots[k] > slots[k - 1]), "dense slots must ascend within 0..119");
        s.slot[k] = slots[k];
        if (slots[k] < 64) s.act_lo |= 1ull << slots[k];
        else s.act_hi |= 1ull << (slots[k] - 64);
    }
    return TDP_OK;
}

static int dn_fill(uint8_t *T, uint64_t n, hipStream_t st) {
    const uint64_t words = n / 4;
    if (words) {
        k_dn_fill<<<dgrid(words), kSpBlock, 0, st>>>(reinterpret_cast<uint32_t *>(T), words);
        TDP_LAUNCH_CHECK();
    }
    if (n & 3) {
        k_dn_fill_tail<<<1, 64, 0, st>>>(T + 4 * words, n & 3);
        TDP_LAUNCH_CHECK();
    }
    return TDP_OK;
}

}  // namespace tdp

using namespace tdp;

extern "C" {

int tdp_dn_from_sparse(const void *d_ents, uint64_t n, const uint8_t *slots, int w, uint8_t *d_T, void *stream) {
    DnSlots s;
    int rc = dn_slots(slots, w, s);
    if (rc) return rc;
    TDP_REQUIRE(((uintptr_t)d_T & 3) == 0, "dense table not 4-byte aligned");
    TDP_REQUIRE(n <= (1ull << w), "%llu states for a dense table of 2^%d", (unsigned long long)n, w);
    hipStream_t st = as_stream(stream);
    rc = dn_fill(d_T, 1ull << w, st);
    if (rc) return rc;
    if (n) {
        k_dn_scatter<<<dgrid(n), kSpBlock, 0, st>>>(static_cast<const Ent *>(d_ents), n, s, d_T);
        TDP_LAUNCH_CHECK();
    }
    return TDP_OK;
}

int tdp_dn_to_sparse(const uint8_t *d_T, int w, const uint8_t *slots, void *d_out, uint64_t cap,
                     unsigned long long *d_counters, void *stream) {
    DnSlots s;
    int rc = dn_slots(slots, w, s);
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    k_dn_set<<<1, 64, 0, st>>>(d_counters, 0ull, 2);
    TDP_LAUNCH_CHECK();
    k_dn_gather<<<dgrid(1ull << w), kSpBlock, 0, st>>>(d_T, 1ull << w, s, static_cast<Ent *>(d_out), cap,
                                                       d_counters);
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

int tdp_dn_step(const uint8_t *d_T, int w_old, uint8_t *d_U, int w_mid, const uint8_t *start_pos, int nstart,
                const uint8_t *end_pos, const uint8_t *end_bit, int nend, uint64_t cm, int cap, int K, int ncls,
                const uint64_t *cls_mask, const int64_t *cls_w, int64_t extra, uint64_t tmask, int use_h,
                unsigned long long *d_count, void *stream) {
    TDP_REQUIRE(w_old >= 0 && w_old <= kDnMaxBits && nstart >= 0 && nend >= 0 && w_mid == w_old + nstart &&
                    w_mid - nend >= 0 && w_mid - nend <= kDnMaxBits && w_mid <= 62,
                "dense step widths: old %d, mid %d, %d starting, %d ending rows", w_old, w_mid, nstart, nend);
    TDP_REQUIRE(cap >= 0 && cap <= 254, "cap %d outside 0..254", cap);
    TDP_REQUIRE(!use_h || (ncls >= 0 && ncls <= kDnMaxCls), "%d lookahead classes (at most %d)", ncls, kDnMaxCls);
    TDP_REQUIRE(!(cm >> w_mid), "update mask outside the mid space");
    DnStepArgs a;
    memset(&a, 0, sizeof(a));
    a.cm = cm;
    a.cap = cap;
    a.K = K;
    a.nstart = nstart;
    a.nend = nend;
    for (int k = 0; k < nstart; k++) {
        TDP_REQUIRE(start_pos[k] < w_mid && (k == 0 || start_pos[k] < start_pos[k - 1]),
                    "starting positions must descend within the mid space");
        a.start_pos[k] = start_pos[k];
        a.start_mask |= 1ull << start_pos[k];
    }
    for (int k = 0; k < nend; k++) {
        TDP_REQUIRE(end_pos[k] < w_mid && (k == 0 || end_pos[k] > end_pos[k - 1]) && end_bit[k] <= 1,
                    "ending positions must ascend within the mid space");
        a.end_pos[k] = end_pos[k];
        a.end_bit[k] = end_bit[k];
    }
    const int w_new = w_mid - nend;
    if (use_h) {
        a.use_h = 1;
        a.extra = extra;
        a.tmask = tmask;
        a.ncls = ncls;
        for (int c = 0; c < ncls; c++) {
            a.cls_mask[c] = cls_mask[c];
            a.cls_w[c] = cls_w[c];
        }
        TDP_REQUIRE(w_new == 64 || !(tmask >> w_new), "target mask outside the new table");
    }
    hipStream_t st = as_stream(stream);
    k_dn_set<<<1, 64, 0, st>>>(d_count, 0ull, 1);
    TDP_LAUNCH_CHECK();
    k_dn_step<<<dgrid(1ull << w_new), kSpBlock, 0, st>>>(d_T, d_U, 1ull << w_new, a, d_count);
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

int tdp_dn_join(const void *d_f, uint64_t nf, const uint8_t *d_T, int w, const uint8_t *slots, const uint64_t *tmask,
                unsigned long long *d_best, void *stream) {
    DnSlots s;
    int rc = dn_slots(slots, w, s);
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    k_dn_set<<<1, 64, 0, st>>>(d_best, ~0ull, 4);
    TDP_LAUNCH_CHECK();
    if (!nf) return TDP_OK;
    const Ent *f = static_cast<const Ent *>(d_f);
    const uint64_t tlo = tmask[0], thi = tmask[1] & kKeyMask;
    k_dn_join_total<<<dgrid(nf), kSpBlock, 0, st>>>(f, nf, d_T, s, tlo, thi, d_best);
    TDP_LAUNCH_CHECK();
    for (int stage = 1; stage <= 3; stage++) {
        k_dn_join_pick<<<dgrid(nf), kSpBlock, 0, st>>>(f, nf, d_T, s, tlo, thi, stage, d_best);
        TDP_LAUNCH_CHECK();
    }
    return TDP_OK;
}

}  // extern "C"
