// Min-plus XOR joins of the tree-realization DP on gfx950 (cdist/tree_dp.c:13-51).
//
// A join enumerates the 2^J compatible (left, right, parent) index triples of one tree node: generators
// 0..pr-1 lift the parent's unit vectors, generators pr..J-1 span the kernel of the child-state sum, and the
// left / right indices are XORs of the generator maps lm / rm.  out[p] = min(out[p], L[l] + R[r]) over the
// triples whose costs are finite and whose sum stays below the infinity.
//
// Full joins: LPP lanes (a power of two <= 64, inside one wave) share a parent p; each lane walks 2^kr / LPP free
// combinations in Gray order (one XOR per step), the lanes reduce their minima with wave shuffles and one lane
// writes out[p] -- no atomics.  Partial Gray-counter ranges [start, stop) (the reference's chunks) use one thread
// per counter and an atomic minimum on the output word.
#include "tdp_common.h"

namespace tdp {

struct JoinMaps {
    uint64_t lm[64];
    uint64_t rm[64];
};

constexpr int kTreeBlock = 256;

template <typename T>
__device__ __forceinline__ void atomic_min_item(T *addr, T v);

template <>
__device__ __forceinline__ void atomic_min_item<uint64_t>(uint64_t *addr, uint64_t v) {
    atomicMin(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)v);
}

template <>
__device__ __forceinline__ void atomic_min_item<uint32_t>(uint32_t *addr, uint32_t v) {
    atomicMin(addr, v);
}

// 8- and 16-bit items: compare-and-swap on the aligned 32-bit word that holds them.
template <typename T>
__device__ __forceinline__ void atomic_min_narrow(T *addr, T v) {
    uintptr_t a = reinterpret_cast<uintptr_t>(addr);
    uint32_t *word = reinterpret_cast<uint32_t *>(a & ~(uintptr_t)3);
    unsigned shift = (unsigned)(a & 3) * 8;
    uint32_t mask = ((1u << (8 * sizeof(T))) - 1) << shift;
    uint32_t old = *word;
    for (;;) {
        uint32_t cur = (old & mask) >> shift;
        if (cur <= v) return;
        uint32_t next = (old & ~mask) | ((uint32_t)v << shift);
        uint32_t seen = atomicCAS(word, old, next);
        if (seen == old) return;
        old = seen;
    }
}

template <>
__device__ __forceinline__ void atomic_min_item<uint16_t>(uint16_t *addr, uint16_t v) { atomic_min_narrow(addr, v); }

template <>
__device__ __forceinline__ void atomic_min_item<uint8_t>(uint8_t *addr, uint8_t v) { atomic_min_narrow(addr, v); }

template <typename T>
__global__ __launch_bounds__(kTreeBlock) void k_tree_join_full(const T *__restrict__ L, const T *__restrict__ R,
                                                               T *__restrict__ out, const JoinMaps maps, int pbits,
                                                               int kbits, int lpp_log, uint64_t inf) {
    uint64_t tid = (uint64_t)blockIdx.x * kTreeBlock + threadIdx.x;
    uint64_t p = tid >> lpp_log;
    unsigned sub = (unsigned)(tid & ((1u << lpp_log) - 1));
    bool active = p < (1ull << pbits);
    uint64_t best = inf;
    if (active) {
        uint64_t lp = 0, rp = 0;
        for (int b = 0; b < pbits; b++)
            if ((p >> b) & 1) { lp ^= maps.lm[b]; rp ^= maps.rm[b]; }
        int per_log = kbits - lpp_log;
        uint64_t per = 1ull << per_log, base = (uint64_t)sub << per_log;
        uint64_t li = lp, ri = rp;
        for (int b = 0; b < kbits; b++)
            if ((base >> b) & 1) { li ^= maps.lm[pbits + b]; ri ^= maps.rm[pbits + b]; }
        for (uint64_t t = 0; t < per; t++) {
            if (t) {
                int bit = __builtin_ctzll(t);
                li ^= maps.lm[pbits + bit];
                ri ^= maps.rm[pbits + bit];
            }
            uint64_t a = L[li], b = R[ri];
            if (a < inf && b < inf && a <= inf - b && a + b < best) best = a + b;
        }
    }
    // minimum over the lanes of this parent inside the wave; with more than 64 lanes per parent each wave's
    // minimum goes to out[p] by an atomic minimum
    int red_log = lpp_log < 6 ? lpp_log : 6;
    for (int off = 1; off < (1 << red_log); off <<= 1) {
        uint64_t o = __shfl_xor(best, off, 64);
        best = o < best ? o : best;
    }
    if (active && (sub & ((1u << red_log) - 1)) == 0) {
        if (lpp_log <= 6) {
            if (best < (uint64_t)out[p]) out[p] = (T)best;
        } else if (best < inf) {
            atomic_min_item<T>(out + p, (T)best);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kTreeBlock) void k_tree_join_range(const T *__restrict__ L, const T *__restrict__ R,
                                                                T *out, const JoinMaps maps, int pbits,
                                                                uint64_t start, uint64_t stop, uint64_t inf) {
    uint64_t i = start + (uint64_t)blockIdx.x * kTreeBlock + threadIdx.x;
    if (i >= stop) return;
    uint64_t g = i ^ (i >> 1), li = 0, ri = 0, pi = 0;
    for (uint64_t rest = g; rest; rest &= rest - 1) {
        int b = __builtin_ctzll(rest);
        li ^= maps.lm[b];
        ri ^= maps.rm[b];
        if (b < pbits) pi |= 1ull << b;
    }
    uint64_t a = L[li], b = R[ri];
    if (a < inf && b < inf && a <= inf - b) atomic_min_item<T>(out + pi, (T)(a + b));
}

template <typename T>
__global__ void k_tree_leaf(T *out, uint64_t size, uint64_t v0, uint64_t v1, uint64_t fill) {
    uint64_t i = (uint64_t)blockIdx.x * kTreeBlock + threadIdx.x;
    if (i < size) out[i] = (T)(i == 0 ? v0 : (i == 1 ? v1 : fill));
}

template <typename T>
__global__ void k_tree_fill(T *out, uint64_t size, uint64_t v) {
    uint64_t i = (uint64_t)blockIdx.x * kTreeBlock + threadIdx.x;
    if (i < size) out[i] = (T)v;
}

__global__ void k_copy_bytes(uint8_t *__restrict__ dst, const uint8_t *__restrict__ src, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * kTreeBlock + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

static unsigned tgrid(uint64_t items) { return (unsigned)((items + kTreeBlock - 1) / kTreeBlock); }

template <typename T>
static int tree_join_t(const void *left, const void *right, void *out, const JoinMaps &maps, unsigned pbits,
                       unsigned gens, uint64_t start, uint64_t stop, uint64_t inf, hipStream_t st) {
    const T *L = static_cast<const T *>(left);
    const T *R = static_cast<const T *>(right);
    T *O = static_cast<T *>(out);
    uint64_t total = 1ull << gens;
    if (start >= stop) return TDP_OK;
    if (start == 0 && stop == total) {
        // lanes per parent: up to one wave, more (with per-wave atomics) when the parents alone cannot fill the
        // chip (about 2^20 threads)
        int kbits = (int)(gens - pbits);
        int lpp_log = kbits < 6 ? kbits : 6;
        if ((int)pbits + lpp_log < 20) lpp_log = kbits < 20 - (int)pbits ? kbits : 20 - (int)pbits;
        uint64_t threads = (1ull << pbits) << lpp_log;
        k_tree_join_full<T><<<tgrid(threads), kTreeBlock, 0, st>>>(L, R, O, maps, (int)pbits, kbits, lpp_log, inf);
    } else {
        k_tree_join_range<T><<<tgrid(stop - start), kTreeBlock, 0, st>>>(L, R, O, maps, (int)pbits, start, stop,
                                                                          inf);
    }
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

}  // namespace tdp

using namespace tdp;

extern "C" {

int tdp_tree_join(int itemsize, const void *d_left, const void *d_right, void *d_out, const uint64_t *lm,
                  const uint64_t *rm, unsigned parent_bits, unsigned generators, uint64_t start, uint64_t stop,
                  uint64_t infinity, void *stream) {
    TDP_REQUIRE(generators <= 62 && parent_bits <= generators, "join with %u generators, %u parent bits",
                generators, parent_bits);
    TDP_REQUIRE(stop <= (1ull << generators), "counter range beyond 2^%u", generators);
    JoinMaps maps;
    memset(&maps, 0, sizeof(maps));
    for (unsigned b = 0; b < generators; b++) {
        maps.lm[b] = lm[b];
        maps.rm[b] = rm[b];
    }
    hipStream_t st = as_stream(stream);
    switch (itemsize) {
        case 1: return tree_join_t<uint8_t>(d_left, d_right, d_out, maps, parent_bits, generators, start, stop, infinity, st);
        case 2: return tree_join_t<uint16_t>(d_left, d_right, d_out, maps, parent_bits, generators, start, stop, infinity, st);
        case 4: return tree_join_t<uint32_t>(d_left, d_right, d_out, maps, parent_bits, generators, start, stop, infinity, st);
        case 8: return tree_join_t<uint64_t>(d_left, d_right, d_out, maps, parent_bits, generators, start, stop, infinity, st);
    }
    TDP_REQUIRE(false, "item size %d", itemsize);
}

#define TDP_TREE_JOIN_NAMED(BITS, BYTES)                                                                     \
    int tdp_tree_join_u##BITS(const void *d_left, const void *d_right, void *d_out, const uint64_t *lm,      \
                              const uint64_t *rm, unsigned parent_bits, unsigned generators, uint64_t start, \
                              uint64_t stop, uint64_t infinity, void *stream) {                            \
        return tdp_tree_join(BYTES, d_left, d_right, d_out, lm, rm, parent_bits, generators, start, stop,     \
                             infinity, stream);                                                             \
    }
TDP_TREE_JOIN_NAMED(8, 1)
TDP_TREE_JOIN_NAMED(16, 2)
TDP_TREE_JOIN_NAMED(32, 4)
TDP_TREE_JOIN_NAMED(64, 8)

int tdp_tree_fill(int itemsize, void *d_out, uint64_t size, uint64_t value, void *stream) {
    hipStream_t st = as_stream(stream);
    if (!size) return TDP_OK;
    switch (itemsize) {
        case 1: k_tree_fill<uint8_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint8_t *)d_out, size, value); break;
        case 2: k_tree_fill<uint16_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint16_t *)d_out, size, value); break;
        case 4: k_tree_fill<uint32_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint32_t *)d_out, size, value); break;
        case 8: k_tree_fill<uint64_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint64_t *)d_out, size, value); break;
        default: TDP_REQUIRE(false, "item size %d", itemsize);
    }
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

int tdp_tree_leaf(int itemsize, void *d_out, uint64_t size, uint64_t v0, uint64_t v1, uint64_t fill, void *stream) {
    hipStream_t st = as_stream(stream);
    TDP_REQUIRE(size >= 1, "empty leaf table");
    switch (itemsize) {
        case 1: k_tree_leaf<uint8_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint8_t *)d_out, size, v0, v1, fill); break;
        case 2: k_tree_leaf<uint16_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint16_t *)d_out, size, v0, v1, fill); break;
        case 4: k_tree_leaf<uint32_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint32_t *)d_out, size, v0, v1, fill); break;
        case 8: k_tree_leaf<uint64_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint64_t *)d_out, size, v0, v1, fill); break;
        default: TDP_REQUIRE(false, "item size %d", itemsize);
    }
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

// Overlap-safe move toward lower addresses (the postorder arena compaction of cdist/tree_dp.c:46-51): copied in
// chunks no longer than the distance between source and destination, in increasing order, so no chunk reads a
// byte an earlier chunk wrote.
int tdp_tree_move(void *d_dst, const void *d_src, size_t bytes, void *stream) {
    hipStream_t st = as_stream(stream);
    uint8_t *dst = static_cast<uint8_t *>(d_dst);
    const uint8_t *src = static_cast<const uint8_t *>(d_src);
    if (!bytes || dst == src) return TDP_OK;
    TDP_REQUIRE(dst < src, "tree_move only moves toward lower addresses");
    uint64_t gap = (uint64_t)(src - dst);
    for (uint64_t off = 0; off < bytes; off += gap) {
        uint64_t len = bytes - off < gap ? bytes - off : gap;
        if (len >= (1u << 16) && gap >= bytes) {
            TDP_HIP_CHECK(hipMemcpyAsync(dst + off, src + off, len, hipMemcpyDeviceToDevice, st));
        } else {
            k_copy_bytes<<<tgrid(len), kTreeBlock, 0, st>>>(dst + off, src + off, len);
            TDP_LAUNCH_CHECK();
        }
    }
    return TDP_OK;
}

}  // extern "C"
