// Batched dense trellis DP with witnesses: many independent plans in one persistent launch (gfx950).
//
// Each workgroup takes the next plan from a device queue (atomic counter; plans sorted by decreasing work on the
// host, so long plans start first) and runs it alone, op by op, with a workgroup barrier between ops: no other
// workgroup touches its table, so plans of any width and length mix freely and no launch is spent per op.
// The table lives in global memory (2^peak bytes, ping-pong for selects); plans up to ~2^21 entries stay in the
// L2/MALL of their XCD.
//
// Witness (when requested): the forward pass copies the table to a checkpoint every S ops; then, segment by segment
// from the last, the workgroup recomputes the segment from its checkpoint writing one decision bit per state and
// update (bit set = the state took its partner's value, after < before), and one thread walks the segment's ops
// backward: select -> re-insert the target bit, expand -> the new top bit must be 0, update -> if the state's bit
// is set the column is used and state ^= mask.  This is the traceback of cdist/auto.py:126-182 with decision bits
// in place of stored tables.
#include "tdp_common.h"

namespace tdp {

constexpr int kBatchThreads = 1024;

struct BatchItem {
    const uint64_t *ops;     // device copy of the op stream (code, a, b)
    int64_t nops;
    int32_t peak_bits;
    int32_t S;               // checkpoint interval (witness only)
    uint8_t *tab0, *tab1;    // 2^peak bytes each
    uint8_t *ckpt;           // ceil(nops / S) tables (witness only)
    uint64_t *ck_n;          // table size at each checkpoint
    uint32_t *bits;          // S decision maps of 2^peak bits (witness only)
    int64_t *used;           // op indices of the used updates (witness only)
    int64_t *result;         // {value, nused, error, traceback state, traceback value, traceback error}
    uint32_t *barrier;       // {arrivals, generation} of the plan's workgroup group
    int32_t witness;
    int32_t group;           // workgroups sharing the plan
};

struct Ticket {
    int32_t item, member;
};

// A member that waits at a group barrier longer than this (s_memrealtime ticks, 100 MHz: 10 s) marks the plan
// failed (error 5); every later barrier of a failed plan returns at once, so the launch always drains.
constexpr uint64_t kBarrierTimeout = 1000000000ull;

// ----------------------------------------------------------------------------- group-wide table operations
// A plan is run by `size` workgroups; work unit u of an op belongs to thread (u mod size * 1024) of the group.
struct Par {
    uint64_t start, stride;
    bool lead;               // member 0, thread 0
};

__device__ __forceinline__ Par par_of(int member, int size) {
    return Par{(uint64_t)member * kBatchThreads + threadIdx.x, (uint64_t)size * kBatchThreads,
               member == 0 && threadIdx.x == 0};
}

// all workgroups of a group: every thread's global writes of this op are visible to the group afterwards
// (the pattern of a grid barrier: workgroup barrier, device-scope release by one thread, arrival counter and
// generation word, device-scope acquire, workgroup barrier)
__device__ void group_sync(uint32_t *bar, int size, uint32_t &gen, int64_t *result) {
    __syncthreads();
    if (size > 1) {
        if (threadIdx.x == 0) {
            unsigned long long *err = reinterpret_cast<unsigned long long *>(&result[2]);
            __threadfence();
            uint32_t g = gen;
            if (atomicAdd(&bar[0], 1u) == (uint32_t)size - 1) {
                atomicExch(&bar[0], 0u);
                __threadfence();
                atomicAdd(&bar[1], 1u);
            } else {
                uint64_t t0 = __builtin_amdgcn_s_memrealtime();
                while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 5ull) break;
                    if (__builtin_amdgcn_s_memrealtime() - t0 > kBarrierTimeout) {
                        atomicExch(err, 5ull);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            gen = g + 1;
            __threadfence();
        }
        __syncthreads();
    }
}
__device__ __forceinline__ uint32_t neq_bytes4(uint32_t a, uint32_t b) {
    // bit i (0..3) set iff byte i of a differs from byte i of b
    uint32_t d = a ^ b;
    d |= d >> 4;
    d |= d >> 2;
    d |= d >> 1;
    d &= 0x01010101u;
    return (d | (d >> 7) | (d >> 14) | (d >> 21)) & 0xFu;
}

// out byte i = in byte (i ^ m) over 32 bytes held as 8 words, m in 0..31
__device__ __forceinline__ void xor_bytes32(const uint32_t *in, uint32_t *out, unsigned m) {
    unsigned q = m >> 2, r = m & 3;
#pragma unroll
    for (int d = 0; d < 8; d++) out[d] = xor_bytes4(in[d ^ q], r);
}

__device__ void wg_update(const Par &P, uint8_t *T, uint64_t n, uint64_t h, uint32_t w, uint32_t *bits) {
    if (h == 0 || n < 2) return;
    if (w > 255) w = 255;
    int hb = 63 - __builtin_clzll(h);
    if (n < 32) {
        if (P.lead) {
            uint32_t word = 0;
            uint64_t half = 1ull << hb;
            for (uint64_t s = 0; s < n; s++) {
                if (s & half) continue;
                uint64_t t = s ^ h;
                uint8_t a = T[s], b = T[t];
                uint8_t na = sat_add1(b, w) < a ? sat_add1(b, w) : a;
                uint8_t nb = sat_add1(a, w) < b ? sat_add1(a, w) : b;
                if (na != a) word |= 1u << s;
                if (nb != b) word |= 1u << t;
                T[s] = na;
                T[t] = nb;
            }
            if (bits) bits[0] = word;
        }
        return;
    }
    uint32_t *W = reinterpret_cast<uint32_t *>(T);
    if (hb >= 5) {
        uint64_t hl = h ^ (1ull << hb);
        uint64_t hlb = hl >> 5;
        unsigned m = (unsigned)(hl & 31);
        int hb5 = hb - 5;
        uint64_t npairs = n >> 6;
        for (uint64_t c = P.start; c < npairs; c += P.stride) {
            uint64_t low = c & ((1ull << hb5) - 1);
            uint64_t xb = ((c >> hb5) << (hb5 + 1)) | low;
            uint64_t yb = (xb ^ hlb) | (1ull << hb5);
            uint32_t x[8], y[8], xp[8], yp[8], bx = 0, by = 0;
#pragma unroll
            for (int d = 0; d < 8; d++) { x[d] = W[8 * xb + d]; y[d] = W[8 * yb + d]; }
            xor_bytes32(x, xp, m);
            xor_bytes32(y, yp, m);
#pragma unroll
            for (int d = 0; d < 8; d++) {
                uint32_t nx = min4(x[d], sat_add4(yp[d], w));
                uint32_t ny = min4(y[d], sat_add4(xp[d], w));
                bx |= neq_bytes4(nx, x[d]) << (4 * d);
                by |= neq_bytes4(ny, y[d]) << (4 * d);
                W[8 * xb + d] = nx;
                W[8 * yb + d] = ny;
            }
            if (bits) { bits[xb] = bx; bits[yb] = by; }
        }
    } else {
        uint64_t nblk = n >> 5;
        for (uint64_t c = P.start; c < nblk; c += P.stride) {
            uint32_t x[8], xp[8], b = 0;
#pragma unroll
            for (int d = 0; d < 8; d++) x[d] = W[8 * c + d];
            xor_bytes32(x, xp, (unsigned)h);
#pragma unroll
            for (int d = 0; d < 8; d++) {
                uint32_t nx = min4(x[d], sat_add4(xp[d], w));
                b |= neq_bytes4(nx, x[d]) << (4 * d);
                W[8 * c + d] = nx;
            }
            if (bits) bits[c] = b;
        }
    }
}

__device__ void wg_fill(const Par &P, uint8_t *T, uint64_t n, uint8_t v) {
    if (n < 4) {
        if (P.start < n) T[P.start] = v;
        return;
    }
    uint32_t *W = reinterpret_cast<uint32_t *>(T);
    for (uint64_t i = P.start; i < (n >> 2); i += P.stride) W[i] = 0x01010101u * v;
}

__device__ void wg_copy(const Par &P, uint8_t *dst, const uint8_t *src, uint64_t n) {
    if (n < 4) {
        if (P.start < n) dst[P.start] = src[P.start];
        return;
    }
    const uint32_t *S = reinterpret_cast<const uint32_t *>(src);
    uint32_t *D = reinterpret_cast<uint32_t *>(dst);
    for (uint64_t i = P.start; i < (n >> 2); i += P.stride) D[i] = S[i];
}

// dst[k] = src[k with bit pos = target inserted], k < n/2
__device__ void wg_select(const Par &P, uint8_t *dst, const uint8_t *src, uint64_t n, int pos, uint32_t target) {
    uint64_t half = n >> 1;
    if (pos >= 2 && half >= 4) {
        const uint32_t *S = reinterpret_cast<const uint32_t *>(src);
        uint32_t *D = reinterpret_cast<uint32_t *>(dst);
        int p2 = pos - 2;
        for (uint64_t k = P.start; k < (half >> 2); k += P.stride) {
            uint64_t low = k & ((1ull << p2) - 1);
            D[k] = S[((k >> p2) << (p2 + 1)) | ((uint64_t)target << p2) | low];
        }
    } else {
        uint64_t lowmask = (1ull << pos) - 1;
        for (uint64_t k = P.start; k < half; k += P.stride)
            dst[k] = src[((k & ~lowmask) << 1) | ((uint64_t)target << pos) | (k & lowmask)];
    }
}

// runs ops [first, last) on the table; bits (when not null) receives one decision map per update, the k-th update
// of the range at bits + k * words
__device__ void wg_run(const Par &P, const BatchItem &B, uint32_t &gen, int64_t first, int64_t last, uint8_t *&cur,
                       uint8_t *&other, uint64_t &n, uint32_t *bits, uint64_t words) {
    int64_t k = 0;
    const uint64_t *ops = B.ops;
    for (int64_t i = first; i < last; i++) {
        uint64_t code = ops[3 * i], a = ops[3 * i + 1], b = ops[3 * i + 2];
        if (code == 0) {
            wg_fill(P, cur + n, n, 255);
            n <<= 1;
        } else if (code == 1) {
            wg_update(P, cur, n, a, (uint32_t)b, bits ? bits + k * words : nullptr);
            k++;
        } else {
            wg_select(P, other, cur, n, (int)a, (uint32_t)b);
            uint8_t *t = cur;
            cur = other;
            other = t;
            n >>= 1;
        }
        group_sync(B.barrier, B.group, gen, B.result);
    }
}

__global__ __launch_bounds__(kBatchThreads) void k_batch(BatchItem *items, const Ticket *tickets, int ntickets,
                                                          int *next) {
    __shared__ int ticket;
    for (;;) {
        if (threadIdx.x == 0) ticket = atomicAdd(next, 1);
        __syncthreads();
        int tk = ticket;
        __syncthreads();
        if (tk >= ntickets) return;
        Ticket T = tickets[tk];
        BatchItem B = items[T.item];
        Par P = par_of(T.member, B.group);
        uint32_t gen = 0;
        int64_t *R = B.result;              // {value, nused, error, state, value, traceback error}
        uint8_t *cur = B.tab0, *other = B.tab1;
        uint64_t n = 1;
        uint64_t cap = 1ull << B.peak_bits;
        if (P.lead) cur[0] = 0;
        group_sync(B.barrier, B.group, gen, R);
        int64_t S = B.witness ? B.S : B.nops + 1;
        for (int64_t g0 = 0; g0 < B.nops; g0 += S) {
            int64_t g1 = g0 + S < B.nops ? g0 + S : B.nops;
            if (B.witness) {
                wg_copy(P, B.ckpt + (g0 / S) * cap, cur, n);
                if (P.lead) B.ck_n[g0 / S] = n;
                group_sync(B.barrier, B.group, gen, R);
            }
            wg_run(P, B, gen, g0, g1, cur, other, n, nullptr, 0);
        }
        if (P.lead) {
            int v = cur[0];
            R[0] = v;
            R[1] = 0;
            if (n != 1) atomicExch(reinterpret_cast<unsigned long long *>(&R[2]), 4ull);
            R[3] = 0;
            R[4] = v;
            R[5] = (n == 1 && v != 255) ? 0 : -1;
        }
        group_sync(B.barrier, B.group, gen, R);
        if (B.witness && R[5] == 0 && B.nops > 0) {
            uint64_t words = cap >= 32 ? cap >> 5 : 1;
            int64_t nseg = (B.nops + S - 1) / S;
            for (int64_t g = nseg - 1; g >= 0; g--) {
                int64_t first = g * S, last = first + S < B.nops ? first + S : B.nops;
                cur = B.tab0;
                other = B.tab1;
                n = B.ck_n[g];
                wg_copy(P, cur, B.ckpt + g * cap, n);
                group_sync(B.barrier, B.group, gen, R);
                wg_run(P, B, gen, first, last, cur, other, n, B.bits, words);
                if (P.lead && R[5] == 0) {
                    // walk the segment back; n is the table size after the segment
                    uint64_t state = (uint64_t)R[3];
                    int64_t value = R[4], nused = R[1], err = 0;
                    int64_t k = 0;
                    for (int64_t i = first; i < last; i++) k += B.ops[3 * i] == 1;
                    for (int64_t i = last - 1; i >= first && !err; i--) {
                        uint64_t code = B.ops[3 * i], a = B.ops[3 * i + 1], b = B.ops[3 * i + 2];
                        if (code == 2) {
                            n <<= 1;
                            uint64_t low = state & ((1ull << a) - 1);
                            state = ((state >> a) << (a + 1)) | (b << a) | low;
                        } else if (code == 0) {
                            n >>= 1;
                            if (state >= n) err = 1;
                        } else {
                            k--;
                            const uint32_t *map = B.bits + k * words;
                            if ((map[state >> 5] >> (state & 31)) & 1u) {
                                B.used[nused++] = i;
                                state ^= a;
                                value -= (int64_t)b;
                            }
                        }
                    }
                    R[3] = (int64_t)state;
                    R[4] = value;
                    R[1] = nused;
                    R[5] = err;
                }
                group_sync(B.barrier, B.group, gen, R);
            }
            if (P.lead) {
                int64_t e = R[5] ? R[5] : ((R[3] != 0 || R[4] != 0) ? 3 : 0);
                if (e) atomicExch(reinterpret_cast<unsigned long long *>(&R[2]), (unsigned long long)e);
            }
        }
        __syncthreads();
    }
}

}  // namespace tdp

using namespace tdp;

extern "C" {

int tdp_batch_run(void *d_items, const void *d_tickets, int ntickets, int *d_next, int workgroups, void *stream) {
    TDP_REQUIRE(ntickets >= 0 && workgroups >= 1, "ntickets %d workgroups %d", ntickets, workgroups);
    hipStream_t st = as_stream(stream);
    if (!ntickets) return TDP_OK;
    k_batch<<<workgroups, kBatchThreads, 0, st>>>(static_cast<BatchItem *>(d_items),
                                                  static_cast<const Ticket *>(d_tickets), ntickets, d_next);
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

int tdp_batch_item_bytes(void) { return (int)sizeof(BatchItem); }

}  // extern "C"
