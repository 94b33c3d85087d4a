// Dense tables of the sparse sweeps' fallback on gfx950 (the Dense table and its column step of
// cdist/sparse_dp.py:221-289 and 680-707).
//
// A dense table holds every state of the active rows: T[i], uint8 (255 = absent), for the compact index i of a state
// (bit k of i = the state's bit at slot slots[k], slots ascending), 2^w entries.  A sweep switches to it where a
// sparse table fills more than a given fraction of its 2^w states and back where it empties; the tables hold exactly
// the states and values the sparse tables would, so values, counts and witnesses do not change.
//
// One column step is one gather pass writing the new table.  Output entry j is placed in the "mid" space (the old
// slots and the rows starting at the column) by inserting the ending rows' target bits; the old values of it and of
// its partner (xor the column's rows) are read -- absent where a starting row's bit is set -- and
// min(T[x], T[x ^ cm] + 1) is kept if it is within the cap and the lookahead bound.  Both index maps (new -> mid,
// mid -> old) are a few runs of consecutive bits, held in registers.  The bound is linear in the bits of j ^ tmask
// (sum_c w_c |u & m_c| = sum_b u_b W_b, W_b = sum of the weights of the classes holding bit b), so it is read from
// per-byte tables in LDS.  When bits 0-3 (or 0-1) of the index pass through both maps, a lane handles sixteen (four)
// consecutive entries with one load per table and one store; the partner's low bits become a byte permutation.
// Counts are reduced per block before one atomic: atomics on one cache line serialise at about 12 ns each on MI355X,
// so the step's blocks are few and large (1024 threads, at most kDnStepGrid of them).
#include "tdp_sparse.h"

namespace tdp {

constexpr int kDnMaxBits = 40;            // the widest dense table (2^40 bytes)
constexpr int kDnMaxCls = 64;             // lookahead classes of one cut with a nonzero compact mask
constexpr int kDnRuns = 8;                // runs of consecutive bits per index map on the fast path
constexpr int kDnKeyBytes = 15;           // bytes of a 120-bit key
constexpr int kDnIdxBytes = 5;            // bytes of a 40-bit compact index
constexpr uint32_t kDnGridCap = 1u << 20;
constexpr uint32_t kDnLutGrid = 2048;     // blocks of the kernels that first build lookup tables in LDS
constexpr int kDnBlock = 1024;            // threads of a step block
constexpr int kDnUnroll = 4;              // groups in flight per thread of a fast step
constexpr uint32_t kDnStepGrid = 256;     // step blocks at most: one per CU (one count atomic each)
constexpr uint64_t kDnChunk = 1ull << 16; // table entries per block of the dense-to-sparse gather (one atomic each)
constexpr int64_t kDnQ = 720720;
static_assert(kSpBlock == 256, "one LDS table entry per thread and byte value");

struct DnSlots {                          // the compact layout of one table
    uint64_t act_lo, act_hi;              // the active slots as a 120-bit mask
    int32_t w, kb0, kb1, pad;             // width; first and last key byte holding an active slot
    uint8_t slot[kDnMaxBits + 8];         // ascending
};

// Generic step (any layout): positions as lists, lookahead classes as masks.
struct DnStepArgs {
    uint64_t cm;                          // the column's rows, mid space
    uint64_t start_mask;                  // the starting rows' bits, mid space
    uint64_t tmask;                       // target row, new compact space
    int64_t extra;
    int32_t nstart, nend, ncls, cap, K, use_h;
    uint8_t start_pos[kDnMaxBits + 8];    // descending mid positions of the starting rows
    uint8_t end_pos[kDnMaxBits + 8];      // ascending mid positions of the ending rows
    uint8_t end_bit[kDnMaxBits + 8];      // their targets
    uint64_t cls_mask[kDnMaxCls];         // new compact space
    int64_t cls_w[kDnMaxCls];
};

// Fast step: both index maps as runs, the bound as per-bit weights.
struct DnFastArgs {
    uint64_t E;                           // the ending rows' target bits, mid space
    uint64_t S;                           // the starting rows' bits, mid space
    uint64_t cm;                          // the column's rows, mid space
    uint64_t tmask;                       // new compact space
    int64_t extra;
    int64_t W[kDnMaxBits];                // lookahead weight of each bit of the new index
    uint32_t run_new[kDnRuns];            // new index -> mid: source bit | target bit << 8 | length << 16
    uint32_t run_old[kDnRuns];            // mid -> old index, the same packing
    int32_t nrun_new, nrun_old, cap, K, use_h, nlut;
};

__device__ __forceinline__ uint64_t insert_bit(uint64_t x, int pos, uint64_t b) {
    const uint64_t low = x & ((1ull << pos) - 1);
    return ((x >> pos) << (pos + 1)) | (b << pos) | low;
}

__device__ __forceinline__ uint64_t remove_bit(uint64_t x, int pos) {
    const uint64_t low = x & ((1ull << pos) - 1);
    return ((x >> (pos + 1)) << pos) | low;
}

__device__ __forceinline__ uint64_t dn_runs(uint64_t x, const uint32_t *runs, int n) {
    uint64_t out = 0;
#pragma unroll
    for (int r = 0; r < kDnRuns; r++) {
        if (r < n) {
            const uint32_t d = runs[r];
            out |= ((x >> (d & 255)) & ((1ull << (d >> 16)) - 1)) << ((d >> 8) & 255);
        }
    }
    return out;
}

// keep v (one byte value) if within the cap and, with the bound total tot, v + ceil(tot / Q) <= K
__device__ __forceinline__ uint32_t dn_prune(uint32_t v, int cap, int K, bool use_h, int64_t tot) {
    if (v == 255u || (int)v > cap) return 255u;
    if (use_h && (int)v + (int)((tot + kDnQ - 1) / kDnQ) > K) return 255u;
    return v;
}

// ---- lookup tables in LDS (every thread of the block calls these)
// compact bits of key byte b holding value v: c[b][v]
__device__ __forceinline__ void dn_lut_compact(const DnSlots &s, uint8_t *sl, uint64_t (*c)[256]) {
    if ((int)threadIdx.x < s.w) sl[threadIdx.x] = s.slot[threadIdx.x];
    for (int b = 0; b < kDnKeyBytes; b++) c[b][threadIdx.x] = 0;
    __syncthreads();
    const uint32_t v = threadIdx.x;
    for (int k = 0; k < s.w; k++) {
        const uint32_t slot = sl[k];
        if ((v >> (slot & 7)) & 1) c[slot >> 3][v] |= 1ull << k;
    }
    __syncthreads();
}

// key bits of compact byte b holding value v: lo[b][v], hi[b][v]
__device__ __forceinline__ void dn_lut_expand(const DnSlots &s, uint8_t *sl, uint64_t (*lo)[256],
                                              uint64_t (*hi)[256]) {
    if ((int)threadIdx.x < s.w) sl[threadIdx.x] = s.slot[threadIdx.x];
    for (int b = 0; b < kDnIdxBytes; b++) lo[b][threadIdx.x] = hi[b][threadIdx.x] = 0;
    __syncthreads();
    const uint32_t v = threadIdx.x;
    for (int k = 0; k < s.w; k++) {
        if ((v >> (k & 7)) & 1) {
            const uint32_t slot = sl[k];
            if (slot < 64) lo[k >> 3][v] |= 1ull << slot;
            else hi[k >> 3][v] |= 1ull << (slot - 64);
        }
    }
    __syncthreads();
}

__device__ __forceinline__ uint64_t dn_compact(const DnSlots &s, const uint64_t (*c)[256], uint64_t lo, uint64_t hi) {
    uint64_t i = 0;
    for (int b = s.kb0; b <= s.kb1; b++) i |= c[b][(uint32_t)((b < 8 ? lo >> (8 * b) : hi >> (8 * (b - 8))) & 255)];
    return i;
}

// value of a state in a dense table (255 if absent or outside the active slots)
__device__ __forceinline__ uint32_t dn_find(const uint8_t *__restrict__ T, const DnSlots &s,
                                            const uint64_t (*c)[256], uint64_t lo, uint64_t hi) {
    if ((lo & ~s.act_lo) | (hi & ~s.act_hi)) return 255u;
    return T[dn_compact(s, c, lo, hi)];
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
    __shared__ uint8_t sl[kDnMaxBits + 8];
    __shared__ uint64_t c[kDnKeyBytes][256];
    dn_lut_compact(s, sl, c);
    const uint64_t stride = (uint64_t)gridDim.x * kSpBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x; i < n; i += stride) {
        const Ent x = ents[i];
        T[dn_compact(s, c, x.lo, x.hv & kKeyMask)] = (uint8_t)(x.hv >> 56);
    }
}

// Block b converts entries [b * kDnChunk, (b + 1) * kDnChunk): it counts the chunk's states, reserves their range
// with one atomic, then writes them in index order (a running offset in LDS, no further atomics).
__global__ __launch_bounds__(kSpBlock) void k_dn_gather(const uint8_t *__restrict__ T, uint64_t n, DnSlots s,
                                                        Ent *__restrict__ out, uint64_t cap,
                                                        unsigned long long *counters) {
    constexpr int kWaves = kSpBlock / 64;
    __shared__ uint8_t sl[kDnMaxBits + 8];
    __shared__ uint64_t xlo[kDnIdxBytes][256], xhi[kDnIdxBytes][256];
    __shared__ uint32_t s_n[kWaves];
    __shared__ unsigned long long s_at;
    dn_lut_expand(s, sl, xlo, xhi);
    const int nb = (s.w + 7) >> 3;
    const uint64_t lo_i = (uint64_t)blockIdx.x * kDnChunk;
    const uint64_t hi_i = lo_i + kDnChunk < n ? lo_i + kDnChunk : n;
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t mine = 0;
    for (uint64_t i = lo_i + threadIdx.x; i < hi_i; i += kSpBlock) mine += T[i] != 255u;
    for (int o = 32; o >= 1; o >>= 1) mine += (uint32_t)__shfl_xor((int)mine, o, 64);
    if (lane == 0) s_n[w] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int q = 0; q < kWaves; q++) tot += s_n[q];
        s_at = tot ? atomicAdd(&counters[0], tot) : 0ull;
    }
    __syncthreads();
    uint64_t at = s_at;
    for (uint64_t base = lo_i; base < hi_i; base += kSpBlock) {
        const uint64_t i = base + threadIdx.x;
        const uint32_t v = i < hi_i ? T[i] : 255u;
        const uint64_t bal = __ballot(v != 255u);
        __syncthreads();                      // s_n of the previous round has been read
        if (lane == 0) s_n[w] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint64_t p = at + (uint64_t)__popcll(bal & lt);
        for (int q = 0; q < kWaves; q++) {
            if (q < (int)w) p += s_n[q];
            at += s_n[q];
        }
        if (v != 255u) {
            uint64_t lo = 0, hi = 0;
            for (int b = 0; b < nb; b++) {
                const uint32_t byte = (uint32_t)(i >> (8 * b)) & 255;
                lo |= xlo[b][byte];
                hi |= xhi[b][byte];
            }
            if (p < cap) out[p] = Ent{lo, hi | ((uint64_t)v << 56)};
            else counters[1] = 1;
        }
    }
}

// Sum of per-thread counts over a kDnBlock block, one atomic add from thread 0 (every thread must call it).
__device__ __forceinline__ void dn_block_count(uint32_t kept, unsigned long long *count) {
    __shared__ uint32_t s_k[kDnBlock / 64];
    for (int o = 32; o >= 1; o >>= 1) kept += (uint32_t)__shfl_xor((int)kept, o, 64);
    if (__lane_id() == 0) s_k[threadIdx.x >> 6] = kept;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int q = 0; q < kDnBlock / 64; q++) tot += s_k[q];
        if (tot) atomicAdd(count, tot);
    }
}

// Generic step, one entry per lane: any number of starting and ending rows.
__device__ __forceinline__ uint32_t dn_mid(const uint8_t *__restrict__ T, const DnStepArgs &a, uint64_t x) {
    if (x & a.start_mask) return 255u;
    for (int k = 0; k < a.nstart; k++) x = remove_bit(x, a.start_pos[k]);
    return T[x];
}

__global__ __launch_bounds__(kDnBlock) void k_dn_step(const uint8_t *__restrict__ T, uint8_t *__restrict__ U,
                                                      uint64_t nout, DnStepArgs a, unsigned long long *count) {
    const uint64_t stride = (uint64_t)gridDim.x * kDnBlock;
    uint32_t kept = 0;
    for (uint64_t j = (uint64_t)blockIdx.x * kDnBlock + threadIdx.x; j < nout; j += stride) {
        uint64_t x = j;
        for (int k = 0; k < a.nend; k++) x = insert_bit(x, a.end_pos[k], a.end_bit[k]);
        uint32_t v = dn_mid(T, a, x);
        if (a.cm) {
            const uint32_t b = dn_mid(T, a, x ^ a.cm);
            const uint32_t b1 = b >= 254u ? 255u : b + 1;
            v = b1 < v ? b1 : v;
        }
        int64_t tot = a.extra;
        if (a.use_h) {
            const uint64_t u = j ^ a.tmask;
            for (int c = 0; c < a.ncls; c++) tot += a.cls_w[c] * (int64_t)__popcll(u & a.cls_mask[c]);
        }
        v = dn_prune(v, a.cap, a.K, a.use_h != 0, tot);
        U[j] = (uint8_t)v;
        kept += v != 255u;
    }
    dn_block_count(kept, count);
}

// Fast step: a lane handles W consecutive entries (16, 4 or 1: as many as the low bits that pass unchanged through
// both index maps allow), with one load per table and one store of W bytes; the partner's low bits become a byte
// permutation inside the loaded bytes.
template <int W>
__device__ __forceinline__ uint4 dn_load(const uint8_t *__restrict__ T, uint64_t y) {
    if (W == 16) return *reinterpret_cast<const uint4 *>(T + y);
    if (W == 4) return make_uint4(*reinterpret_cast<const uint32_t *>(T + y), 0u, 0u, 0u);
    return make_uint4(T[y], 0u, 0u, 0u);
}

template <int W>
__global__ __launch_bounds__(kDnBlock) void k_dn_step_fast(const uint8_t *__restrict__ T, uint8_t *__restrict__ U,
                                                           uint64_t ngroups, DnFastArgs a,
                                                           unsigned long long *count) {
    constexpr int R = kDnUnroll;                        // groups in flight per thread
    constexpr uint64_t low = W - 1;
    __shared__ int64_t lut[kDnIdxBytes][256];
    if (a.use_h) {
        if (threadIdx.x < 256) {
            for (int k = 0; k < a.nlut; k++) {
                int64_t s = 0;
#pragma unroll
                for (int b = 0; b < 8; b++)
                    if ((threadIdx.x >> b) & 1) s += a.W[8 * k + b];
                lut[k][threadIdx.x] = s;
            }
        }
        __syncthreads();
    }
    const uint32_t none = W == 1 ? 255u : 0xFFFFFFFFu;
    const uint64_t stride = (uint64_t)gridDim.x * kDnBlock;
    uint32_t kept = 0;
    for (uint64_t g0 = (uint64_t)blockIdx.x * kDnBlock + threadIdx.x; g0 < ngroups; g0 += R * stride) {
        uint4 v[R], b[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint64_t g = g0 + r * stride;
            v[r] = b[r] = make_uint4(none, none, none, none);
            if (g >= ngroups) continue;
            const uint64_t j = g * W;
            const uint64_t x = dn_runs(j, a.run_new, a.nrun_new) | a.E;
            const uint64_t xp = x ^ (a.cm & ~low);
            if (!(x & a.S)) v[r] = dn_load<W>(T, dn_runs(x, a.run_old, a.nrun_old));
            if (a.cm && !(xp & a.S)) b[r] = dn_load<W>(T, dn_runs(xp, a.run_old, a.nrun_old));
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint64_t g = g0 + r * stride;
            if (g >= ngroups) continue;
            const uint64_t j = g * W;
            int64_t hi = a.extra;
            const uint64_t u = j ^ a.tmask;
            if (a.use_h)
                for (int k = 1; k < a.nlut; k++) hi += lut[k][(u >> (8 * k)) & 255];
            uint4 w = v[r];
            if (a.cm) {
                if (W == 16) {
                    w = min16(w, sat_add16(xor_bytes16(b[r], (unsigned)(a.cm & 15)), 1u));
                } else if (W == 4) {
                    w.x = min4(w.x, sat_add4(xor_bytes4(b[r].x, (unsigned)(a.cm & 3)), 1u));
                } else {
                    const uint32_t b1 = b[r].x >= 254u ? 255u : b[r].x + 1;
                    w.x = b1 < w.x ? b1 : w.x;
                }
            }
            uint32_t o[4] = {0u, 0u, 0u, 0u};
            const uint32_t in[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int i = 0; i < W; i++) {
                const int64_t tot = a.use_h && a.nlut ? hi + lut[0][(u ^ i) & 255] : hi;
                const uint32_t e = dn_prune((in[i >> 2] >> (8 * (i & 3))) & 255u, a.cap, a.K, a.use_h != 0, tot);
                o[i >> 2] |= e << (8 * (i & 3));
                kept += e != 255u;
            }
            if (W == 16) *reinterpret_cast<uint4 *>(U + j) = make_uint4(o[0], o[1], o[2], o[3]);
            else if (W == 4) *reinterpret_cast<uint32_t *>(U + j) = o[0];
            else U[j] = (uint8_t)o[0];
        }
    }
    dn_block_count(kept, count);
}

__global__ __launch_bounds__(kSpBlock) void k_dn_join_total(const Ent *__restrict__ f, uint64_t nf,
                                                            const uint8_t *__restrict__ T, DnSlots s, uint64_t tm_lo,
                                                            uint64_t tm_hi, unsigned long long *best) {
    __shared__ uint8_t sl[kDnMaxBits + 8];
    __shared__ uint64_t c[kDnKeyBytes][256];
    dn_lut_compact(s, sl, c);
    const uint64_t stride = (uint64_t)gridDim.x * kSpBlock;
    uint32_t t = ~0u;
    for (uint64_t i = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x; i < nf; i += stride) {
        const Ent x = f[i];
        const uint32_t vb = dn_find(T, s, c, x.lo ^ tm_lo, (x.hv & kKeyMask) ^ tm_hi);
        if (vb != 255u) t = min(t, (uint32_t)(x.hv >> 56) + vb);
    }
    block_min_atomic(t, &best[0]);
}

// stage 1: the least key hi among states reaching the least total; 2: then the least key lo; 3: its value
__global__ __launch_bounds__(kSpBlock) void k_dn_join_pick(const Ent *__restrict__ f, uint64_t nf,
                                                           const uint8_t *__restrict__ T, DnSlots s, uint64_t tm_lo,
                                                           uint64_t tm_hi, int stage, unsigned long long *best) {
    __shared__ uint8_t sl[kDnMaxBits + 8];
    __shared__ uint64_t c[kDnKeyBytes][256];
    if (best[0] == ~0ull) return;                     // no join total (the whole block returns): nothing to pick
    dn_lut_compact(s, sl, c);
    const uint64_t stride = (uint64_t)gridDim.x * kSpBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kSpBlock + threadIdx.x; i < nf; i += stride) {
        const Ent x = f[i];
        const uint64_t hi = x.hv & kKeyMask;
        if (stage >= 2 && hi != best[1]) continue;
        if (stage >= 3 && x.lo != best[2]) continue;
        const uint32_t vb = dn_find(T, s, c, x.lo ^ tm_lo, hi ^ tm_hi);
        if (vb == 255u || (x.hv >> 56) + vb != best[0]) continue;
        if (stage == 1) atomicMin(&best[1], (unsigned long long)hi);
        else if (stage == 2) atomicMin(&best[2], (unsigned long long)x.lo);
        else best[3] = x.hv >> 56;
    }
}

__global__ void k_dn_set(unsigned long long *p, unsigned long long v, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

static unsigned dgrid(uint64_t items, uint32_t cap = kDnGridCap, int block = kSpBlock) {
    const uint64_t g = (items + block - 1) / block;
    return (unsigned)(g == 0 ? 1 : (g > cap ? cap : g));
}

static int dn_slots(const uint8_t *slots, int w, DnSlots &s) {
    TDP_REQUIRE(w >= 0 && w <= kDnMaxBits, "dense table of %d bits (at most %d)", w, kDnMaxBits);
    memset(&s, 0, sizeof(s));
    s.w = w;
    s.kb0 = 0;
    s.kb1 = -1;
    for (int k = 0; k < w; k++) {
        TDP_REQUIRE(slots[k] < 120 && (k == 0 || slots[k] > slots[k - 1]), "dense slots must ascend within 0..119");
        s.slot[k] = slots[k];
        if (slots[k] < 64) s.act_lo |= 1ull << slots[k];
        else s.act_hi |= 1ull << (slots[k] - 64);
    }
    if (w) {
        s.kb0 = slots[0] >> 3;
        s.kb1 = slots[w - 1] >> 3;
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

// Runs of consecutive bits of the positions below `width` outside `skip`, in order: bit p of the source (mid) goes
// to bit k of the target, k counting the kept positions (to_mid: the reverse direction, kept k -> mid p).
static int dn_make_runs(uint64_t skip, int width, bool to_mid, uint32_t *runs) {
    int n = 0, k = 0, last = -2;
    for (int p = 0; p < width; p++) {
        if ((skip >> p) & 1) continue;
        if (p == last + 1 && n) {
            runs[n - 1] += 1u << 16;
        } else {
            if (n == kDnRuns) return -1;
            runs[n++] = to_mid ? ((uint32_t)k | ((uint32_t)p << 8) | (1u << 16))
                               : ((uint32_t)p | ((uint32_t)k << 8) | (1u << 16));
        }
        last = p;
        k++;
    }
    return n;
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
        k_dn_scatter<<<dgrid(n, kDnLutGrid), kSpBlock, 0, st>>>(static_cast<const Ent *>(d_ents), n, s, d_T);
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
    const uint64_t n = 1ull << w;
    k_dn_gather<<<(unsigned)((n + kDnChunk - 1) / kDnChunk), kSpBlock, 0, st>>>(d_T, n, s, static_cast<Ent *>(d_out),
                                                                               cap, d_counters);
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
    TDP_REQUIRE(((uintptr_t)d_T & 3) == 0 && ((uintptr_t)d_U & 3) == 0, "dense tables not 4-byte aligned");
    const int w_new = w_mid - nend;
    uint64_t S = 0, Emask = 0, E = 0;
    for (int k = 0; k < nstart; k++) {
        TDP_REQUIRE(start_pos[k] < w_mid && (k == 0 || start_pos[k] < start_pos[k - 1]),
                    "starting positions must descend within the mid space");
        S |= 1ull << start_pos[k];
    }
    for (int k = 0; k < nend; k++) {
        TDP_REQUIRE(end_pos[k] < w_mid && (k == 0 || end_pos[k] > end_pos[k - 1]) && end_bit[k] <= 1,
                    "ending positions must ascend within the mid space");
        Emask |= 1ull << end_pos[k];
        E |= (uint64_t)end_bit[k] << end_pos[k];
    }
    TDP_REQUIRE(!use_h || !(tmask >> w_new), "target mask outside the new table");
    hipStream_t st = as_stream(stream);
    k_dn_set<<<1, 64, 0, st>>>(d_count, 0ull, 1);
    TDP_LAUNCH_CHECK();

    DnFastArgs f;
    memset(&f, 0, sizeof(f));
    f.nrun_new = dn_make_runs(Emask, w_mid, true, f.run_new);
    f.nrun_old = dn_make_runs(S, w_mid, false, f.run_old);
    if (f.nrun_new >= 0 && f.nrun_old >= 0) {
        f.E = E;
        f.S = S;
        f.cm = cm;
        f.cap = cap;
        f.K = K;
        if (use_h) {
            f.use_h = 1;
            f.extra = extra;
            f.tmask = tmask;
            f.nlut = (w_new + 7) / 8;
            for (int c = 0; c < ncls; c++)
                for (int b = 0; b < w_new; b++)
                    if ((cls_mask[c] >> b) & 1) f.W[b] += cls_w[c];
        }
        // sixteen (four) entries per lane when bits 0-3 (0-1) of the new index are those of the mid and old indexes
        if (w_new >= 4 && !((S | Emask) & 15)) {
            const uint64_t groups = 1ull << (w_new - 4);
            k_dn_step_fast<16><<<dgrid(groups, kDnStepGrid, kDnBlock), kDnBlock, 0, st>>>(d_T, d_U, groups, f,
                                                                                         d_count);
        } else if (w_new >= 2 && !((S | Emask) & 3)) {
            const uint64_t groups = 1ull << (w_new - 2);
            k_dn_step_fast<4><<<dgrid(groups, kDnStepGrid, kDnBlock), kDnBlock, 0, st>>>(d_T, d_U, groups, f,
                                                                                        d_count);
        } else {
            k_dn_step_fast<1><<<dgrid(1ull << w_new, kDnStepGrid, kDnBlock), kDnBlock, 0, st>>>(
                d_T, d_U, 1ull << w_new, f, d_count);
        }
        TDP_LAUNCH_CHECK();
        return TDP_OK;
    }
    DnStepArgs a;
    memset(&a, 0, sizeof(a));
    a.cm = cm;
    a.start_mask = S;
    a.cap = cap;
    a.K = K;
    a.nstart = nstart;
    a.nend = nend;
    for (int k = 0; k < nstart; k++) a.start_pos[k] = start_pos[k];
    for (int k = 0; k < nend; k++) {
        a.end_pos[k] = end_pos[k];
        a.end_bit[k] = end_bit[k];
    }
    if (use_h) {
        a.use_h = 1;
        a.extra = extra;
        a.tmask = tmask;
        a.ncls = ncls;
        for (int c = 0; c < ncls; c++) {
            a.cls_mask[c] = cls_mask[c];
            a.cls_w[c] = cls_w[c];
        }
    }
    k_dn_step<<<dgrid(1ull << w_new, kDnStepGrid, kDnBlock), kDnBlock, 0, st>>>(d_T, d_U, 1ull << w_new, a, d_count);
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
    k_dn_join_total<<<dgrid(nf, kDnStepGrid), kSpBlock, 0, st>>>(f, nf, d_T, s, tlo, thi, d_best);
    TDP_LAUNCH_CHECK();
    for (int stage = 1; stage <= 3; stage++) {
        k_dn_join_pick<<<dgrid(nf, kDnLutGrid), kSpBlock, 0, st>>>(f, nf, d_T, s, tlo, thi, stage, d_best);
        TDP_LAUNCH_CHECK();
    }
    return TDP_OK;
}

}  // extern "C"
