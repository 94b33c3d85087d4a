// Meet-in-the-middle lookup tables on gfx950: all column subsets of size <= k keyed by syndrome hash (the tables of
// cdist/gmitm.c:36-186, built and probed on the device).
//
// Open addressing over 2^log2slots slots, linear probing from zh & mask; key 0 = empty.
//   compact (<= 16 observables): key = bit63 | bits 16..62 of zl (fingerprint) | 16-bit logical label;
//   wide:                        key = bit63 | zl, the full label in lgs[], cost 255 until the label is published.
// Insert: CAS the key into an empty slot; a slot holding the same key keeps the least cost (byte minimum by CAS
// on its 32-bit word).  A wide insert that meets a matching slot whose label is not yet published does not wait
// (lanes of one wave cannot wait on each other): it is written to a retry list and inserted again by the next
// launch, after the owner has published.
// Build: one thread per leading index pair (i < j) walks the remaining indices depth first, so every subset of
// size 2..k is inserted once; singletons and the empty set come from their own launch.
#include "tdp_tables.h"

#include <mutex>
#include <new>
#include <vector>

namespace tdp {

constexpr int kTabBlock = 256;
constexpr int kMaxK = 8;

struct Retry {
    uint64_t zh, zl, lg;
    uint32_t cost;
};

__device__ __forceinline__ void cost_min(uint8_t *cost, uint64_t i, uint8_t v) {
    uintptr_t a = reinterpret_cast<uintptr_t>(cost + i);
    uint32_t *word = reinterpret_cast<uint32_t *>(a & ~(uintptr_t)3);
    unsigned shift = (unsigned)(a & 3) * 8;
    uint32_t old = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        uint32_t cur = (old >> shift) & 0xFFu;
        if (cur <= v) return;
        uint32_t next = (old & ~(0xFFu << shift)) | ((uint32_t)v << shift);
        uint32_t seen = atomicCAS(word, old, next);
        if (seen == old) return;
        old = seen;
    }
}

__device__ __forceinline__ uint8_t cost_load(const uint8_t *cost, uint64_t i) {
    uintptr_t a = reinterpret_cast<uintptr_t>(cost + i);
    const uint32_t *word = reinterpret_cast<const uint32_t *>(a & ~(uintptr_t)3);
    uint32_t w = __hip_atomic_load(const_cast<uint32_t *>(word), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    return (uint8_t)((w >> ((a & 3) * 8)) & 0xFFu);
}

// returns false when a wide insert must be retried by a later launch; `added` counts the keys this thread stored
// (the caller adds its total to the table count once, not one atomic per key on a counter every thread shares)
__device__ bool tab_insert(const TabView &t, bool wide, uint64_t zh, uint64_t zl, uint64_t lg, uint8_t c,
                           uint32_t &added) {
    uint64_t i = zh & t.mask;
    if (!wide) {
        uint64_t key = (zl & kFpMask) | (lg & 0xFFFF) | (1ull << 63);
        for (;;) {
            uint64_t cur = __hip_atomic_load(&t.key[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cur == 0) {
                uint64_t seen = atomicCAS(reinterpret_cast<unsigned long long *>(&t.key[i]), 0ull,
                                          (unsigned long long)key);
                if (seen == 0) {
                    added++;
                    cur = key;
                } else {
                    cur = seen;
                }
            }
            if (cur == key) {
                cost_min(t.cost, i, c);
                return true;
            }
            i = (i + 1) & t.mask;
        }
    }
    uint64_t key = zl | (1ull << 63);
    for (;;) {
        uint64_t cur = __hip_atomic_load(&t.key[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == 0) {
            uint64_t seen = atomicCAS(reinterpret_cast<unsigned long long *>(&t.key[i]), 0ull,
                                      (unsigned long long)key);
            if (seen == 0) {
                added++;
                __hip_atomic_store(&t.lgs[i], lg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __threadfence();
                cost_min(t.cost, i, c);
                return true;
            }
            cur = seen;
        }
        if (cur == key) {
            if (cost_load(t.cost, i) == 255) return false;           // owner has not published: retry later
            __threadfence();
            if (__hip_atomic_load(&t.lgs[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == lg) {
                cost_min(t.cost, i, c);
                return true;
            }
        }
        i = (i + 1) & t.mask;
    }
}

__device__ __forceinline__ void insert_or_defer(const TabView &t, bool wide, uint64_t zh, uint64_t zl, uint64_t lg,
                                                uint8_t c, Retry *retry, unsigned long long *nretry,
                                                uint64_t retry_cap, uint32_t &added) {
    if (!tab_insert(t, wide, zh, zl, lg, c, added)) {
        unsigned long long r = atomicAdd(nretry, 1ull);
        if (r < retry_cap) retry[r] = Retry{zh, zl, lg, c};
    }
}

// Adds the block's sum of v to *dst with one atomic (every thread of the block calls it): atomics on one cache line
// serialise at about 12 ns each, so one per thread of a launch would take longer than the build.
__device__ __forceinline__ void block_count_add(uint32_t v, unsigned long long *dst) {
    __shared__ uint32_t s_sum;
    if (threadIdx.x == 0) s_sum = 0;
    __syncthreads();
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&s_sum, v);
    __syncthreads();
    if (threadIdx.x == 0 && s_sum) atomicAdd(dst, (unsigned long long)s_sum);
}

__global__ void k_tab_clear(TabView t, uint64_t n, int wide) {
    uint64_t i = (uint64_t)blockIdx.x * kTabBlock + threadIdx.x;
    if (i >= n) return;
    t.key[i] = 0;
    t.cost[i] = 255;
    if (wide) t.lgs[i] = 0;
}

// the empty set (cost 0) and the singletons
__global__ void k_tab_small(TabView t, int wide, int n, const uint64_t *zh, const uint64_t *zl, const uint64_t *lg,
                            Retry *retry, unsigned long long *nretry, uint64_t retry_cap) {
    int j = (int)(blockIdx.x * kTabBlock + threadIdx.x) - 1;
    uint32_t added = 0;
    if (j < 0) insert_or_defer(t, wide, 0, 0, 0, 0, retry, nretry, retry_cap, added);
    else if (j < n) insert_or_defer(t, wide, zh[j], zl[j], lg[j], 1, retry, nretry, retry_cap, added);
    block_count_add(added, t.count);
}

// subsets of size 2..k with leading pair (i, j), i < j, remaining indices ascending (depth first)
// (one thread's share: returns the number of entries it added to the table)
__device__ uint32_t tab_pairs_one(uint64_t idx, TabView t, int wide, int n, int k, const uint64_t *zh,
                                  const uint64_t *zl, const uint64_t *lg, Retry *retry, unsigned long long *nretry,
                                  uint64_t retry_cap) {
    uint64_t npairs = (uint64_t)n * (uint64_t)(n - 1) / 2;
    if (idx >= npairs) return 0;
    // unrank idx -> (i, j) with i < j: row i holds n - 1 - i pairs
    int i = 0;
    uint64_t rem = idx;
    {
        // closed form start, then correct
        double nn = (double)n;
        double disc = (2 * nn - 1) * (2 * nn - 1) - 8.0 * (double)idx;
        int guess = (int)((2 * nn - 1 - sqrt(disc > 0 ? disc : 0)) / 2);
        if (guess < 0) guess = 0;
        if (guess > n - 2) guess = n - 2;
        i = guess;
        uint64_t before = (uint64_t)i * (uint64_t)(2 * n - i - 1) / 2;
        while (i > 0 && before > idx) {
            i--;
            before = (uint64_t)i * (uint64_t)(2 * n - i - 1) / 2;
        }
        while (i < n - 2 && (uint64_t)(i + 1) * (uint64_t)(2 * n - i - 2) / 2 <= idx) {
            i++;
            before = (uint64_t)i * (uint64_t)(2 * n - i - 1) / 2;
        }
        rem = idx - before;
    }
    int j = i + 1 + (int)rem;
    uint64_t h[kMaxK + 1], l[kMaxK + 1], g[kMaxK + 1];
    int at[kMaxK + 1];
    h[2] = zh[i] ^ zh[j];
    l[2] = zl[i] ^ zl[j];
    g[2] = lg[i] ^ lg[j];
    uint32_t added = 0;
    insert_or_defer(t, wide, h[2], l[2], g[2], 2, retry, nretry, retry_cap, added);
    int depth = 2;           // subset size of h[depth]
    at[depth] = j;           // last index used
    int next = j + 1;
    while (k > 2) {
        if (next < n) {
            int d1 = depth + 1;
            h[d1] = h[depth] ^ zh[next];
            l[d1] = l[depth] ^ zl[next];
            g[d1] = g[depth] ^ lg[next];
            insert_or_defer(t, wide, h[d1], l[d1], g[d1], (uint8_t)d1, retry, nretry, retry_cap, added);
            if (d1 < k) {
                at[d1] = next;
                depth = d1;
                next = next + 1;
            } else {
                next++;
            }
        } else {
            if (depth == 2) break;
            next = at[depth] + 1;
            depth--;
        }
    }
    return added;
}

__global__ __launch_bounds__(kTabBlock) void k_tab_pairs(TabView t, int wide, int n, int k, const uint64_t *zh,
                                                         const uint64_t *zl, const uint64_t *lg, Retry *retry,
                                                         unsigned long long *nretry, uint64_t retry_cap) {
    const uint64_t idx = (uint64_t)blockIdx.x * kTabBlock + threadIdx.x;
    block_count_add(tab_pairs_one(idx, t, wide, n, k, zh, zl, lg, retry, nretry, retry_cap), t.count);
}

__global__ void k_tab_retry(TabView t, const Retry *in, uint64_t m, Retry *out, unsigned long long *nout,
                            uint64_t cap) {
    uint64_t i = (uint64_t)blockIdx.x * kTabBlock + threadIdx.x;
    uint32_t added = 0;
    if (i < m) {
        Retry r = in[i];
        insert_or_defer(t, true, r.zh, r.zl, r.lg, (uint8_t)r.cost, out, nout, cap, added);
    }
    block_count_add(added, t.count);
}

// Least cost <= R of an entry with the query's fingerprint and a different label; -1 if none (gmitm.c:136-163).
__global__ void k_tab_lookup(TabView t, int wide, const uint64_t *zh, const uint64_t *zl, const uint64_t *lg,
                             const int *R, uint64_t m, int *out) {
    uint64_t q = (uint64_t)blockIdx.x * kTabBlock + threadIdx.x;
    if (q >= m) return;
    out[q] = tab_probe(t, wide, zh[q], zl[q], lg[q], R[q]);
}

__global__ void k_zero_words(unsigned long long *w, int n) {
    if ((int)threadIdx.x < n) w[threadIdx.x] = 0;
}

static unsigned tab_grid(uint64_t n) { return (unsigned)((n + kTabBlock - 1) / kTabBlock); }


// Shared body of the table builds: columns copied to d_cols (3 x max(n, 1) words), retry lists rbuf (2 x retry_cap
// records) and their two counters rcnt, all device memory the caller provides.  Null stream, synchronous.
static int build_table(DevTable *t, int n, const uint64_t *zh, const uint64_t *zl, const uint64_t *lg, int k,
                       uint64_t *d_cols, Retry *rbuf, unsigned long long *rcnt, uint64_t retry_cap) {
    t->k = k;
    const size_t cols = n > 0 ? (size_t)n : 1;
    uint64_t *dzh = d_cols, *dzl = d_cols + cols, *dlg = d_cols + 2 * cols;
    if (n) {
        TDP_HIP_CHECK(hipMemcpy(dzh, zh, n * 8, hipMemcpyHostToDevice));
        TDP_HIP_CHECK(hipMemcpy(dzl, zl, n * 8, hipMemcpyHostToDevice));
        TDP_HIP_CHECK(hipMemcpy(dlg, lg, n * 8, hipMemcpyHostToDevice));
    }
    k_zero_words<<<1, 64>>>(rcnt, 2);
    TabView v = view(t);
    int rc = TDP_OK;
    int small = k >= 1 ? n + 1 : 1;
    k_tab_small<<<tab_grid(small), kTabBlock>>>(v, t->wide, k >= 1 ? n : 0, dzh, dzl, dlg, rbuf, rcnt, retry_cap);
    if (k >= 2 && n >= 2) {
        uint64_t npairs = (uint64_t)n * (uint64_t)(n - 1) / 2;
        k_tab_pairs<<<tab_grid(npairs), kTabBlock>>>(v, t->wide, n, k, dzh, dzl, dlg, rbuf, rcnt, retry_cap);
    }
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        set_error("table build launch failed");
        rc = TDP_EHIP;
    }
    // wide tables: inserts that met an unpublished slot, again until none is left
    int which = 0;
    for (int round = 0; rc == TDP_OK && round < 1000; round++) {
        unsigned long long m = 0;
        if (hipMemcpy(&m, rcnt + which, 8, hipMemcpyDeviceToHost) != hipSuccess) { rc = TDP_EHIP; break; }
        if (m == 0) break;
        if (m > retry_cap) {
            set_error("table build: %llu deferred inserts exceed the retry list of %llu", m,
                      (unsigned long long)retry_cap);
            rc = TDP_ECAPACITY;
            break;
        }
        int other = which ^ 1;
        k_zero_words<<<1, 64>>>(rcnt + other, 1);
        k_tab_retry<<<tab_grid(m), kTabBlock>>>(v, rbuf + which * retry_cap, m, rbuf + other * retry_cap,
                                                rcnt + other, retry_cap);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) { rc = TDP_EHIP; break; }
        which = other;
    }
    if (rc == TDP_EHIP) set_error("table build failed on the device");
    return rc;
}

// ---------------------------------------------------------------------------------------------- static scratch
// Device memory of tdp_tab_build_noalloc: a file-scope array each device holds for the life of the process.  Host
// code never passes the array's name to a kernel (on the host it is only a stand-in): scratch_base() looks up the
// device address once per device, with that device current, and every use goes through that address.
constexpr int kScratchCols = 4096;
constexpr uint64_t kScratchRetry = 1u << 16;
constexpr uint64_t kScratchColBytes = 3ull * kScratchCols * 8;
constexpr uint64_t kScratchBytes = kScratchColBytes + 256 + 2 * kScratchRetry * sizeof(Retry);
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint8_t *g_scratch_ptr[kMaxDevices];

static int scratch_base(uint8_t **out) {
    int dev = 0;
    TDP_HIP_CHECK(hipGetDevice(&dev));
    TDP_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d the static scratch table tracks", dev,
                kMaxDevices);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        TDP_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        TDP_REQUIRE(ptr != nullptr, "no device address for the static scratch array");
        g_scratch_ptr[dev] = static_cast<uint8_t *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return TDP_OK;
}

}  // namespace tdp

using namespace tdp;

extern "C" {

void *tdp_tab_create(int log2slots, int wide) {
    if (log2slots < 1 || log2slots > 40) {
        set_error("log2slots %d out of range", log2slots);
        return nullptr;
    }
    DevTable *t = new (std::nothrow) DevTable();
    if (!t) return nullptr;
    uint64_t n = 1ull << log2slots;
    t->mask = n - 1;
    t->wide = wide;
    t->log2slots = (uint64_t)log2slots;
    bool ok = hipMalloc(&t->key, n * 8) == hipSuccess && hipMalloc(&t->cost, n + 4) == hipSuccess &&
              hipMalloc(&t->count, 8) == hipSuccess && (!wide || hipMalloc(&t->lgs, n * 8) == hipSuccess);
    if (!ok) {
        set_error("device allocation of a 2^%d-slot table failed", log2slots);
        hipFree(t->key);
        hipFree(t->cost);
        hipFree(t->count);
        hipFree(t->lgs);
        delete t;
        return nullptr;
    }
    k_tab_clear<<<tab_grid(n), kTabBlock>>>(view(t), n, wide);
    hipMemset(t->count, 0, 8);
    if (hipDeviceSynchronize() != hipSuccess) {
        set_error("table initialisation failed");
        return nullptr;
    }
    return t;
}

void tdp_tab_free(void *p) {
    DevTable *t = static_cast<DevTable *>(p);
    if (!t) return;
    hipFree(t->key);
    hipFree(t->cost);
    hipFree(t->count);
    hipFree(t->lgs);
    delete t;
}

int64_t tdp_tab_count(void *p) {
    DevTable *t = static_cast<DevTable *>(p);
    unsigned long long c = 0;
    if (!t || hipMemcpy(&c, t->count, 8, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return (int64_t)c;
}

int tdp_tab_build(void *p, int n, const uint64_t *zh, const uint64_t *zl, const uint64_t *lg, int k) {
    DevTable *t = static_cast<DevTable *>(p);
    TDP_REQUIRE(t && n >= 0 && k >= 0 && k <= kMaxK, "tab_build: n %d k %d (k <= %d)", n, k, kMaxK);
    uint64_t *d = nullptr;
    size_t cols = n > 0 ? (size_t)n : 1;
    TDP_HIP_CHECK(hipMalloc(&d, 3 * cols * 8));
    const uint64_t retry_cap = 1u << 20;
    Retry *rbuf = nullptr;
    unsigned long long *rcnt = nullptr;
    int rc;
    if (hipMalloc(&rbuf, 2 * retry_cap * sizeof(Retry)) != hipSuccess ||
        hipMalloc(&rcnt, 2 * sizeof(unsigned long long)) != hipSuccess) {
        set_error("device allocation of the table build's retry lists failed");
        rc = TDP_EHIP;
    } else {
        rc = build_table(t, n, zh, zl, lg, k, d, rbuf, rcnt, retry_cap);
    }
    hipFree(rbuf);
    hipFree(rcnt);
    hipFree(d);
    return rc;
}

int tdp_tab_build_noalloc(void *p, int n, const uint64_t *zh, const uint64_t *zl, const uint64_t *lg, int k) {
    DevTable *t = static_cast<DevTable *>(p);
    TDP_REQUIRE(t && n >= 0 && k >= 0 && k <= kMaxK, "tab_build_noalloc: n %d k %d (k <= %d)", n, k, kMaxK);
    TDP_REQUIRE(n <= kScratchCols, "tab_build_noalloc: %d columns exceed the static scratch limit of %d", n,
                kScratchCols);
    uint8_t *base = nullptr;
    int rc = scratch_base(&base);
    if (rc != TDP_OK) return rc;
    // layout of the scratch array: columns (zh, zl, lg), the retry counters, the two retry lists
    uint64_t *cols = reinterpret_cast<uint64_t *>(base);
    unsigned long long *rcnt = reinterpret_cast<unsigned long long *>(base + kScratchColBytes);
    Retry *rbuf = reinterpret_cast<Retry *>(base + kScratchColBytes + 256);
    return build_table(t, n, zh, zl, lg, k, cols, rbuf, rcnt, kScratchRetry);
}

int tdp_tab_lookup_batch(void *p, const uint64_t *d_zh, const uint64_t *d_zl, const uint64_t *d_lg, const int *d_R,
                         uint64_t m, int *d_out, void *stream) {
    DevTable *t = static_cast<DevTable *>(p);
    TDP_REQUIRE(t != nullptr, "null table");
    if (!m) return TDP_OK;
    k_tab_lookup<<<tab_grid(m), kTabBlock, 0, as_stream(stream)>>>(view(t), t->wide, d_zh, d_zl, d_lg, d_R, m, d_out);
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

}  // extern "C"
